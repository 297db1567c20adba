// Data ingest and egress of llr_recon_exec (host code, no device work):
//   - BART .cfl/.hdr volumes (the cfl crate's read_to_buffer / get_dims /
//     write_buffer used by data_import.rs:37-185 and lib.rs:354-362): the .hdr
//     is text, "# Dimensions" then up to 16 sizes; the .cfl is the raw
//     column-major complex64 data, little endian;
//   - DtiDataSetBuilder (data_import.rs:13-273): the sorted volume files matching
//     a glob pattern, the b-table (.txt, n x [b, gx, gy, gz]), the b0 mask
//     (data_import.rs:322-328), the (ny, nz, nq) phase-encode sample mask, the
//     DTI / b0 volume sets and their masks;
//   - LlrReconParams read from / written to TOML (lib.rs:47-80,
//     bin/llr-recon.rs:6-25, bin/llr-reco-params.rs:14-23);
//   - error_history.json (lib.rs:345-351).
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include <glob.h>

#include "../../include/llr_hip.h"

namespace llr {
llr_status fail(llr_status st, const std::string& msg);
}

using llr::fail;

namespace {

// path without a trailing .cfl / .hdr, then with `ext` (Path::with_extension)
std::string with_ext(const std::string& p, const char* ext) {
  std::string base = p;
  const size_t slash = base.find_last_of('/');
  const size_t dot = base.find_last_of('.');
  if (dot != std::string::npos && (slash == std::string::npos || dot > slash)) base = base.substr(0, dot);
  return base + ext;
}

llr_status read_hdr(const std::string& path, std::vector<size_t>& dims) {
  std::ifstream f(with_ext(path, ".hdr"));
  if (!f) return fail(LLR_ERR_ARG, "failed to open header " + with_ext(path, ".hdr"));
  std::string line;
  dims.clear();
  while (std::getline(f, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream is(line);
    long long d;
    while (is >> d) {
      if (d < 1) return fail(LLR_ERR_ARG, "non-positive dimension in " + with_ext(path, ".hdr"));
      dims.push_back((size_t)d);
    }
    break;
  }
  if (dims.empty() || dims.size() > 16) return fail(LLR_ERR_ARG, "no dimensions in " + with_ext(path, ".hdr"));
  return LLR_OK;
}

size_t prod(const std::vector<size_t>& d) {
  size_t n = 1;
  for (size_t v : d) n *= v;
  return n;
}

std::vector<size_t> non_singleton(const std::vector<size_t>& d) {
  std::vector<size_t> out;
  for (size_t v : d)
    if (v != 1) out.push_back(v);
  return out;
}

llr_status read_cfl_data(const std::string& path, llr_cf32* out, size_t len) {
  std::vector<size_t> dims;
  llr_status st = read_hdr(path, dims);
  if (st) return st;
  if (prod(dims) != len)
    return fail(LLR_ERR_ARG, "buffer of " + std::to_string(len) + " samples for " + with_ext(path, ".cfl") +
                                 " holding " + std::to_string(prod(dims)));
  FILE* f = std::fopen(with_ext(path, ".cfl").c_str(), "rb");
  if (!f) return fail(LLR_ERR_ARG, "failed to open " + with_ext(path, ".cfl"));
  const size_t got = std::fread(out, sizeof(llr_cf32), len, f);
  std::fclose(f);
  if (got != len) return fail(LLR_ERR_ARG, "short read of " + with_ext(path, ".cfl"));
  return LLR_OK;
}

// b0 mask (data_import.rs:322-328): |g| <= tolerance
void b0_mask_of(const float* bvecs, size_t n, float tol, std::vector<int>& m) {
  m.resize(n);
  for (size_t i = 0; i < n; ++i) {
    const float* g = bvecs + 3 * i;
    m[i] = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]) <= tol ? 1 : 0;
  }
}

// shortest decimal that reads back to the same double (serde_json's output)
std::string shortest(double v) {
  if (std::isnan(v) || std::isinf(v)) return "null";  // serde_json writes non-finite floats as null
  char buf[64];
  for (int p = 1; p <= 17; ++p) {
    std::snprintf(buf, sizeof buf, "%.*g", p, v);
    if (std::strtod(buf, nullptr) == v) break;
  }
  std::string s = buf;
  // serde_json prints a float with a fraction or exponent: 1 -> 1.0, 1e-5 -> 1e-5
  if (s.find_first_of(".en") == std::string::npos) s += ".0";
  // "1e-05" -> "1e-5" (Rust's exponent form has no padding or '+')
  const size_t e = s.find('e');
  if (e != std::string::npos) {
    std::string mant = s.substr(0, e), ex = s.substr(e + 1);
    bool neg = false;
    if (!ex.empty() && (ex[0] == '+' || ex[0] == '-')) {
      neg = ex[0] == '-';
      ex = ex.substr(1);
    }
    while (ex.size() > 1 && ex[0] == '0') ex = ex.substr(1);
    s = mant + "e" + (neg ? "-" : "") + ex;
  }
  return s;
}

// the same for an f32 (Rust's Display of f32: shortest digits that read back to
// the same float)
std::string shortest_f32(float v) {
  if (std::isnan(v) || std::isinf(v)) return std::isnan(v) ? "nan" : (v > 0 ? "inf" : "-inf");
  char buf[64];
  for (int p = 1; p <= 9; ++p) {
    std::snprintf(buf, sizeof buf, "%.*g", p, (double)v);
    if (std::strtof(buf, nullptr) == v) break;
  }
  std::string s = buf;
  if (s.find_first_of(".e") == std::string::npos) s += ".0";
  return s;
}

// -------------------------------------------------------------- minimal TOML
// key = value lines (integers, floats, strings, booleans, arrays of numbers),
// '#' comments; enough for LlrReconParams as toml::to_string writes it.
struct TomlValue {
  enum Kind { Num, Str, Arr, Bool } kind = Num;
  double num = 0;
  bool is_int = true;
  std::string str;
  std::vector<double> arr;
};

std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
  return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

bool parse_number(const std::string& t, double* v, bool* is_int) {
  std::string s;
  for (char c : t)
    if (c != '_') s += c;
  if (s.empty()) return false;
  char* end = nullptr;
  errno = 0;
  *v = std::strtod(s.c_str(), &end);
  if (end == s.c_str() || *end != '\0' || errno == ERANGE) return false;
  *is_int = s.find_first_of(".eE") == std::string::npos;
  return true;
}

llr_status parse_toml(const std::string& text, std::vector<std::pair<std::string, TomlValue>>& out,
                      const std::string& path) {
  std::istringstream in(text);
  std::string line;
  int ln = 0;
  while (std::getline(in, line)) {
    ++ln;
    // strip a comment outside a string
    bool inq = false;
    for (size_t i = 0; i < line.size(); ++i) {
      if (line[i] == '"') inq = !inq;
      if (line[i] == '#' && !inq) {
        line = line.substr(0, i);
        break;
      }
    }
    line = trim(line);
    if (line.empty()) continue;
    if (line[0] == '[') return fail(LLR_ERR_ARG, path + ":" + std::to_string(ln) + ": tables are not expected");
    const size_t eq = line.find('=');
    if (eq == std::string::npos) return fail(LLR_ERR_ARG, path + ":" + std::to_string(ln) + ": expected key = value");
    std::string key = trim(line.substr(0, eq)), val = trim(line.substr(eq + 1));
    if (key.size() >= 2 && key.front() == '"' && key.back() == '"') key = key.substr(1, key.size() - 2);
    TomlValue v;
    const std::string where = path + ":" + std::to_string(ln) + ": ";
    if (val.empty()) return fail(LLR_ERR_ARG, where + "missing value for " + key);
    if (val[0] == '"') {
      if (val.size() < 2 || val.back() != '"') return fail(LLR_ERR_ARG, where + "unterminated string");
      v.kind = TomlValue::Str;
      const std::string raw = val.substr(1, val.size() - 2);
      for (size_t i = 0; i < raw.size(); ++i) {
        if (raw[i] == '\\' && i + 1 < raw.size()) {
          const char c = raw[++i];
          v.str += c == 'n' ? '\n' : c == 't' ? '\t' : c;
        } else {
          v.str += raw[i];
        }
      }
    } else if (val[0] == '[') {
      if (val.back() != ']') return fail(LLR_ERR_ARG, where + "arrays must close on their line");
      v.kind = TomlValue::Arr;
      std::string body = val.substr(1, val.size() - 2);
      std::istringstream is(body);
      std::string item;
      while (std::getline(is, item, ',')) {
        item = trim(item);
        if (item.empty()) continue;
        double d;
        bool ii;
        if (!parse_number(item, &d, &ii)) return fail(LLR_ERR_ARG, where + "bad array element '" + item + "'");
        v.arr.push_back(d);
        v.is_int = v.is_int && ii;
      }
    } else if (val == "true" || val == "false") {
      v.kind = TomlValue::Bool;
      v.num = val == "true" ? 1 : 0;
    } else if (!parse_number(val, &v.num, &v.is_int)) {
      return fail(LLR_ERR_ARG, where + "bad value '" + val + "' for " + key);
    }
    for (auto& kv : out)
      if (kv.first == key) return fail(LLR_ERR_ARG, where + "duplicate key " + key);
    out.push_back({key, v});
  }
  return LLR_OK;
}

llr_status get_uint(const TomlValue& v, const std::string& key, size_t* out) {
  if (v.kind != TomlValue::Num || !v.is_int || v.num < 0)
    return fail(LLR_ERR_ARG, key + ": expected a non-negative integer");
  *out = (size_t)v.num;
  return LLR_OK;
}

llr_status get_uint3(const TomlValue& v, const std::string& key, size_t out[3]) {
  if (v.kind != TomlValue::Arr || v.arr.size() != 3 || !v.is_int)
    return fail(LLR_ERR_ARG, key + ": expected an array of 3 integers");
  for (int a = 0; a < 3; ++a) {
    if (v.arr[a] < 0) return fail(LLR_ERR_ARG, key + ": negative entry");
    out[a] = (size_t)v.arr[a];
  }
  return LLR_OK;
}

}  // namespace

struct llr_dataset {
  std::vector<std::string> files;  // sorted volume files (.cfl)
  size_t vol[3] = {0, 0, 0};
  std::vector<float> bvals, bvecs;  // per file
  std::vector<int> b0;              // 1 = b0 volume
  bool have_btable = false;
  std::vector<llr_cf32> mask;       // (ny, nz, n_files), column-major
  bool have_mask = false;

  std::vector<size_t> indices(int which) const {
    std::vector<size_t> idx;
    for (size_t i = 0; i < files.size(); ++i)
      if (which == LLR_VOLUMES_ALL || (which == LLR_VOLUMES_B0) == (b0[i] != 0)) idx.push_back(i);
    return idx;
  }
};

extern "C" {

llr_status llr_cfl_dims(const char* path, size_t dims[16], size_t* ndims) {
  if (!path || !dims || !ndims) return fail(LLR_ERR_ARG, "null argument");
  std::vector<size_t> d;
  llr_status st = read_hdr(path, d);
  if (st) return st;
  for (size_t i = 0; i < 16; ++i) dims[i] = i < d.size() ? d[i] : 1;
  *ndims = d.size();
  return LLR_OK;
}

llr_status llr_cfl_read(const char* path, llr_cf32* data, size_t len) {
  if (!path || (!data && len)) return fail(LLR_ERR_ARG, "null argument");
  return read_cfl_data(path, data, len);
}

llr_status llr_cfl_write(const char* path, const size_t* dims, size_t ndims, const llr_cf32* data, size_t len) {
  if (!path || !dims || (!data && len)) return fail(LLR_ERR_ARG, "null argument");
  if (ndims == 0 || ndims > 16) return fail(LLR_ERR_ARG, "1 to 16 dimensions");
  size_t n = 1;
  for (size_t i = 0; i < ndims; ++i) n *= dims[i];
  if (n != len) return fail(LLR_ERR_ARG, "data length does not match the dimensions");
  {
    std::ofstream h(with_ext(path, ".hdr"));
    if (!h) return fail(LLR_ERR_ARG, "failed to create " + with_ext(path, ".hdr"));
    h << "# Dimensions\n";
    for (size_t i = 0; i < 16; ++i) h << (i < ndims ? dims[i] : 1) << " ";
    h << "\n";
    if (!h) return fail(LLR_ERR_ARG, "failed to write " + with_ext(path, ".hdr"));
  }
  FILE* f = std::fopen(with_ext(path, ".cfl").c_str(), "wb");
  if (!f) return fail(LLR_ERR_ARG, "failed to create " + with_ext(path, ".cfl"));
  const size_t put = len ? std::fwrite(data, sizeof(llr_cf32), len, f) : 0;
  const int rc = std::fclose(f);
  if (put != len || rc != 0) return fail(LLR_ERR_ARG, "failed to write " + with_ext(path, ".cfl"));
  return LLR_OK;
}

llr_status llr_read_b_table(const char* path, float* b_vals, float* b_vecs, size_t capacity, size_t* n) {
  if (!path || !n) return fail(LLR_ERR_ARG, "null argument");
  std::ifstream f(with_ext(path, ".txt"));
  if (!f) return fail(LLR_ERR_ARG, "failed to open b table " + with_ext(path, ".txt"));
  std::vector<float> v;
  std::string tok;
  while (f >> tok) {
    char* end = nullptr;
    const float x = std::strtof(tok.c_str(), &end);
    if (end == tok.c_str() || *end != '\0') return fail(LLR_ERR_ARG, "failed to parse b-table at character: " + tok);
    v.push_back(x);
  }
  if (v.size() % 4) return fail(LLR_ERR_ARG, "number b-table entries must be divisible by 4: [b-value, gx, gy, gz]");
  const size_t nv = v.size() / 4;
  *n = nv;
  if (!b_vals && !b_vecs) return LLR_OK;  // size query
  if (capacity < nv) return fail(LLR_ERR_ARG, "b-table holds " + std::to_string(nv) + " entries, capacity " +
                                                  std::to_string(capacity));
  for (size_t i = 0; i < nv; ++i) {
    if (b_vals) b_vals[i] = v[4 * i];
    if (b_vecs)
      for (int a = 0; a < 3; ++a) b_vecs[3 * i + a] = v[4 * i + 1 + a];
  }
  return LLR_OK;
}

llr_status llr_b0_mask(const float* b_vecs, size_t n, float tolerance, int* mask) {
  if ((!b_vecs || !mask) && n) return fail(LLR_ERR_ARG, "null argument");
  if (!(tolerance >= 0.f)) return fail(LLR_ERR_ARG, "vector magnitude tolerance must be non-negative");
  std::vector<int> m;
  b0_mask_of(b_vecs, n, tolerance, m);
  std::copy(m.begin(), m.end(), mask);
  return LLR_OK;
}

llr_status llr_dataset_from_cfl_dir(const char* dir, const char* pattern, size_t expected, llr_dataset** out) {
  if (!dir || !pattern || !out) return fail(LLR_ERR_ARG, "null argument");
  std::string full = std::string(dir);
  if (!full.empty() && full.back() != '/') full += '/';
  full = with_ext(full + pattern, ".cfl");
  glob_t g;
  std::memset(&g, 0, sizeof g);
  const int rc = glob(full.c_str(), 0, nullptr, &g);
  std::vector<std::string> files;
  if (rc == 0)
    for (size_t i = 0; i < g.gl_pathc; ++i) files.push_back(g.gl_pathv[i]);
  globfree(&g);
  if (rc != 0 && rc != GLOB_NOMATCH) return fail(LLR_ERR_ARG, "failed to read glob pattern " + full);
  if (expected && files.size() != expected)
    return fail(LLR_ERR_ARG, "unexpected number of files found: " + std::to_string(files.size()) + ", expected " +
                                 std::to_string(expected));
  if (files.empty()) return fail(LLR_ERR_ARG, "no files found for " + full);
  std::sort(files.begin(), files.end());  // ascending, assuming numerical naming
  std::vector<size_t> d0;
  llr_status st = read_hdr(files[0], d0);
  if (st) return st;
  d0 = non_singleton(d0);
  for (size_t i = 1; i < files.size(); ++i) {
    std::vector<size_t> d;
    if ((st = read_hdr(files[i], d))) return st;
    if (non_singleton(d) != d0) return fail(LLR_ERR_ARG, "mismatch between first volume dims and volume " + files[i]);
  }
  if (d0.size() != 3) return fail(LLR_ERR_ARG, "expecting n volume dims to be 3");
  llr_dataset* ds = new llr_dataset();
  ds->files = files;
  for (int a = 0; a < 3; ++a) ds->vol[a] = d0[a];
  ds->b0.assign(files.size(), 0);
  *out = ds;
  return LLR_OK;
}

void llr_dataset_destroy(llr_dataset* ds) { delete ds; }

llr_status llr_dataset_with_b_table(llr_dataset* ds, const char* path, float b0_tol) {
  if (!ds || !path) return fail(LLR_ERR_ARG, "null argument");
  size_t n = 0;
  llr_status st = llr_read_b_table(path, nullptr, nullptr, 0, &n);
  if (st) return st;
  if (n != ds->files.size())
    return fail(LLR_ERR_ARG, "expected " + std::to_string(ds->files.size()) + " bvecs, received " + std::to_string(n));
  ds->bvals.assign(n, 0.f);
  ds->bvecs.assign(3 * n, 0.f);
  if ((st = llr_read_b_table(path, ds->bvals.data(), ds->bvecs.data(), n, &n))) return st;
  if (!(b0_tol >= 0.f)) return fail(LLR_ERR_ARG, "vector magnitude tolerance must be non-negative");
  b0_mask_of(ds->bvecs.data(), n, b0_tol, ds->b0);
  ds->have_btable = true;
  return LLR_OK;
}

llr_status llr_dataset_with_sample_mask(llr_dataset* ds, const char* path) {
  if (!ds || !path) return fail(LLR_ERR_ARG, "null argument");
  std::vector<size_t> d;
  llr_status st = read_hdr(path, d);
  if (st) return st;
  const std::vector<size_t> nd = non_singleton(d);
  // a (ny, nz, nq) volume; a trailing singleton nq = 1 keeps its place
  std::vector<size_t> m3 = nd;
  if (m3.size() == 2 && ds->files.size() == 1) m3.push_back(1);
  if (m3.size() != 3) return fail(LLR_ERR_ARG, "sample mask must be a 3-D (ny, nz, nq) array");
  if (m3[2] != ds->files.size()) return fail(LLR_ERR_ARG, "sample mask volumes do not match the data set");
  if (m3[0] != ds->vol[1] || m3[1] != ds->vol[2]) return fail(LLR_ERR_ARG, "phase encoding dimension mismatch");
  ds->mask.assign(prod(d), llr_cf32{0.f, 0.f});
  if ((st = read_cfl_data(path, ds->mask.data(), ds->mask.size()))) return st;
  ds->have_mask = true;
  return LLR_OK;
}

llr_status llr_dataset_info(const llr_dataset* ds, size_t vol[3], size_t* n_volumes, size_t* n_dti, size_t* n_b0) {
  if (!ds) return fail(LLR_ERR_ARG, "null argument");
  if (vol)
    for (int a = 0; a < 3; ++a) vol[a] = ds->vol[a];
  if (n_volumes) *n_volumes = ds->files.size();
  size_t nb0 = 0;
  for (int b : ds->b0) nb0 += b ? 1 : 0;
  if (n_dti) *n_dti = ds->files.size() - nb0;
  if (n_b0) *n_b0 = nb0;
  return LLR_OK;
}

llr_status llr_dataset_file(const llr_dataset* ds, size_t index, char* out, size_t cap) {
  if (!ds || !out) return fail(LLR_ERR_ARG, "null argument");
  if (index >= ds->files.size()) return fail(LLR_ERR_ARG, "no such volume");
  if (ds->files[index].size() + 1 > cap) return fail(LLR_ERR_ARG, "name buffer too small");
  std::memcpy(out, ds->files[index].c_str(), ds->files[index].size() + 1);
  return LLR_OK;
}

llr_status llr_dataset_indices(const llr_dataset* ds, int which, size_t* out, size_t cap, size_t* n) {
  if (!ds || !n) return fail(LLR_ERR_ARG, "null argument");
  if (which != LLR_VOLUMES_ALL && !ds->have_btable) return fail(LLR_ERR_ARG, "you must load a b-table first");
  const std::vector<size_t> idx = ds->indices(which);
  *n = idx.size();
  if (!out) return LLR_OK;
  if (cap < idx.size()) return fail(LLR_ERR_ARG, "index buffer too small");
  std::copy(idx.begin(), idx.end(), out);
  return LLR_OK;
}

llr_status llr_dataset_b_table(const llr_dataset* ds, float* b_vals, float* b_vecs) {
  if (!ds) return fail(LLR_ERR_ARG, "null argument");
  if (!ds->have_btable) return fail(LLR_ERR_ARG, "you must load a b-table first");
  if (b_vals) std::copy(ds->bvals.begin(), ds->bvals.end(), b_vals);
  if (b_vecs) std::copy(ds->bvecs.begin(), ds->bvecs.end(), b_vecs);
  return LLR_OK;
}

llr_status llr_dataset_load(const llr_dataset* ds, int which, llr_cf32* out, size_t len) {
  if (!ds || !out) return fail(LLR_ERR_ARG, "null argument");
  if (which != LLR_VOLUMES_ALL && !ds->have_btable) return fail(LLR_ERR_ARG, "you must load a b-table first");
  const std::vector<size_t> idx = ds->indices(which);
  const size_t N = ds->vol[0] * ds->vol[1] * ds->vol[2];
  if (len != N * idx.size())
    return fail(LLR_ERR_ARG, "buffer must hold " + std::to_string(idx.size()) + " volumes of " + std::to_string(N));
  for (size_t k = 0; k < idx.size(); ++k) {
    llr_status st = read_cfl_data(ds->files[idx[k]], out + k * N, N);
    if (st) return st;
  }
  return LLR_OK;
}

llr_status llr_dataset_sample_mask(const llr_dataset* ds, int which, llr_cf32* out, size_t len) {
  if (!ds || !out) return fail(LLR_ERR_ARG, "null argument");
  if (!ds->have_mask) return fail(LLR_ERR_ARG, "you must load a sample mask first");
  if (which != LLR_VOLUMES_ALL && !ds->have_btable) return fail(LLR_ERR_ARG, "you must load a b-table first");
  const std::vector<size_t> idx = ds->indices(which);
  const size_t pe = ds->vol[1] * ds->vol[2];
  if (len != pe * idx.size()) return fail(LLR_ERR_ARG, "mask buffer must hold ny*nz per selected volume");
  for (size_t k = 0; k < idx.size(); ++k) std::copy_n(ds->mask.data() + idx[k] * pe, pe, out + k * pe);
  return LLR_OK;
}

void llr_recon_params_default(llr_toml_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof *p);
  for (int a = 0; a < 3; ++a) p->llr_block_size[a] = 32;
  p->fft_max_batch_size = 200;
  p->cas_mat_max_batch_size = 200;
  p->llr_block_rank = 6;
  p->wavelet_shrink_factor = 0.005f;
  p->n_iter = 60;
  p->has_subspace = 1;
  std::snprintf(p->subspace_projection_mat, sizeof p->subspace_projection_mat, "%s", "/path/to/matrix.cfl");
  p->llr_mode = LLR_MODE_RANK;
  p->lambda_factor = 0.02;
  p->precision = LLR_PREC_FP32;
  p->storage = LLR_STORE_FP32;
  p->seed = 0;
}

llr_status llr_recon_params_read(const char* path, llr_toml_params* p) {
  if (!path || !p) return fail(LLR_ERR_ARG, "null argument");
  const std::string file = with_ext(path, ".toml");
  std::ifstream f(file);
  if (!f) return fail(LLR_ERR_ARG, "failed to open config file: " + file);
  std::stringstream ss;
  ss << f.rdbuf();
  std::vector<std::pair<std::string, TomlValue>> kv;
  llr_status st = parse_toml(ss.str(), kv, file);
  if (st) return st;
  llr_toml_params q;
  llr_recon_params_default(&q);
  q.has_subspace = 0;
  q.subspace_projection_mat[0] = '\0';
  const char* required[] = {"llr_block_size", "fft_max_batch_size", "cas_mat_max_batch_size", "llr_block_rank",
                            "wavelet_shrink_factor", "n_iter"};
  for (const char* r : required) {
    bool found = false;
    for (auto& e : kv) found = found || e.first == r;
    if (!found) return fail(LLR_ERR_ARG, "failed to parse recon params file with error: missing field `" +
                                             std::string(r) + "`");
  }
  for (auto& e : kv) {
    const std::string& k = e.first;
    const TomlValue& v = e.second;
    if (k == "llr_block_size") {
      if ((st = get_uint3(v, k, q.llr_block_size))) return st;
    } else if (k == "max_rand_shift") {
      if ((st = get_uint3(v, k, q.max_rand_shift))) return st;
      q.has_max_rand_shift = 1;
    } else if (k == "fft_max_batch_size") {
      if ((st = get_uint(v, k, &q.fft_max_batch_size))) return st;
    } else if (k == "cas_mat_max_batch_size") {
      if ((st = get_uint(v, k, &q.cas_mat_max_batch_size))) return st;
    } else if (k == "llr_block_rank") {
      if ((st = get_uint(v, k, &q.llr_block_rank))) return st;
    } else if (k == "wavelet_shrink_factor") {
      if (v.kind != TomlValue::Num) return fail(LLR_ERR_ARG, k + ": expected a number");
      q.wavelet_shrink_factor = (float)v.num;
    } else if (k == "n_iter") {
      if ((st = get_uint(v, k, &q.n_iter))) return st;
    } else if (k == "subspace_projection_mat") {
      if (v.kind != TomlValue::Str) return fail(LLR_ERR_ARG, k + ": expected a string");
      if (v.str.size() + 1 > sizeof q.subspace_projection_mat) return fail(LLR_ERR_ARG, k + ": path too long");
      std::snprintf(q.subspace_projection_mat, sizeof q.subspace_projection_mat, "%s", v.str.c_str());
      q.has_subspace = 1;
    } else if (k == "llr_mode") {  // extensions of this library
      if (v.kind != TomlValue::Str || (v.str != "rank" && v.str != "svt"))
        return fail(LLR_ERR_ARG, k + ": expected \"rank\" or \"svt\"");
      q.llr_mode = v.str == "svt" ? LLR_MODE_SVT : LLR_MODE_RANK;
    } else if (k == "svt_lambda_factor") {
      if (v.kind != TomlValue::Num || !(v.num > 0)) return fail(LLR_ERR_ARG, k + ": expected a positive number");
      q.lambda_factor = v.num;
    } else if (k == "precision") {
      if (v.kind != TomlValue::Str || (v.str != "fp32" && v.str != "mixed"))
        return fail(LLR_ERR_ARG, k + ": expected \"fp32\" or \"mixed\"");
      q.precision = v.str == "mixed" ? LLR_PREC_MIXED : LLR_PREC_FP32;
    } else if (k == "storage") {
      if (v.kind != TomlValue::Str || (v.str != "fp32" && v.str != "fp16"))
        return fail(LLR_ERR_ARG, k + ": expected \"fp32\" or \"fp16\"");
      q.storage = v.str == "fp16" ? LLR_STORE_FP16 : LLR_STORE_FP32;
    } else if (k == "shift_seed") {
      size_t s = 0;
      if ((st = get_uint(v, k, &s))) return st;
      q.seed = (unsigned long long)s;
    }
    // other keys are ignored, as serde does for fields it does not know
  }
  *p = q;
  return LLR_OK;
}

llr_status llr_recon_params_write(const char* path, const llr_toml_params* p) {
  if (!path || !p) return fail(LLR_ERR_ARG, "null argument");
  const std::string file = with_ext(path, ".toml");
  std::ofstream f(file);
  if (!f) return fail(LLR_ERR_ARG, "Can't create file " + file);
  // the fields of LlrReconParams in declaration order, None options left out
  f << "llr_block_size = [" << p->llr_block_size[0] << ", " << p->llr_block_size[1] << ", " << p->llr_block_size[2]
    << "]\n";
  if (p->has_max_rand_shift)
    f << "max_rand_shift = [" << p->max_rand_shift[0] << ", " << p->max_rand_shift[1] << ", " << p->max_rand_shift[2]
      << "]\n";
  f << "fft_max_batch_size = " << p->fft_max_batch_size << "\n";
  f << "cas_mat_max_batch_size = " << p->cas_mat_max_batch_size << "\n";
  f << "llr_block_rank = " << p->llr_block_rank << "\n";
  f << "wavelet_shrink_factor = " << shortest_f32(p->wavelet_shrink_factor) << "\n";
  f << "n_iter = " << p->n_iter << "\n";
  if (p->has_subspace) f << "subspace_projection_mat = \"" << p->subspace_projection_mat << "\"\n";
  f << "\n";
  if (!f) return fail(LLR_ERR_ARG, "Can't write file " + file);
  return LLR_OK;
}

llr_status llr_read_affine_translation(const char* const* files, size_t n, float multiplier, float* out) {
  if ((!files || !out) && n) return fail(LLR_ERR_ARG, "null argument");
  for (size_t i = 0; i < n; ++i) {
    if (!files[i]) return fail(LLR_ERR_ARG, "null file name");
    std::ifstream f(files[i]);
    if (!f) return fail(LLR_ERR_ARG, std::string("failed to open file ") + files[i]);
    std::string line;
    bool found = false;
    while (!found && std::getline(f, line)) {
      if (line.find("Parameters:") == std::string::npos) continue;
      std::vector<float> v;
      std::istringstream is(line);
      std::string tok;
      while (is >> tok) {  // the numeric tokens of the line (split_ascii_whitespace + parse::<f32>)
        char* end = nullptr;
        const float x = std::strtof(tok.c_str(), &end);
        if (end != tok.c_str() && *end == '\0') v.push_back(x);
      }
      if (v.size() != 12) return fail(LLR_ERR_ARG, std::string("expected 12 matrix entries in ") + files[i]);
      for (int a = 0; a < 3; ++a) out[3 * i + a] = v[9 + a] * multiplier;
      found = true;
    }
    if (!found) return fail(LLR_ERR_ARG, std::string("failed to find Parameter field in affine file ") + files[i]);
  }
  return LLR_OK;
}

llr_status llr_write_error_history(const char* path, const double* hist, size_t n) {
  if (!path || (!hist && n)) return fail(LLR_ERR_ARG, "null argument");
  std::ofstream f(path);
  if (!f) return fail(LLR_ERR_ARG, "failed to create json file " + std::string(path));
  // serde_json::to_string_pretty of a Vec<f64>, plus the newline lib.rs:348 appends
  if (n == 0) {
    f << "[]\n";
  } else {
    f << "[\n";
    for (size_t i = 0; i < n; ++i) f << "  " << shortest(hist[i]) << (i + 1 < n ? ",\n" : "\n");
    f << "]\n";
  }
  if (!f) return fail(LLR_ERR_ARG, "failed to write to json file " + std::string(path));
  return LLR_OK;
}

}  // extern "C"
