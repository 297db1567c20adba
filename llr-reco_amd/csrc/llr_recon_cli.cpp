// llr-recon: the command-line reconstruction of bin/llr-recon.rs:6-25 and
// llr_recon_exec (lib.rs:113-363) on one GPU, through the library's C ABI.
//
//   llr-recon <recon_params> <input_cfl_dir> <input_cfl_pattern> <output_cfl_dir>
//             <output_cfl_prefix> <b_table> [b0_null_vec_tol] [sample_mask_cfl]
//             [expected_volumes] [is_image] [recon_b0]
//             [--seed N] [--device N]
//
// The positional arguments are DataSetParams (lib.rs:83-110) in their order;
// recon_params names a TOML file of LlrReconParams (lib.rs:47-80). Steps, as
// llr_recon_exec: volume files -> b-table -> optional sample mask -> the DTI
// (or b0) set -> optional forward model of image data -> undersampling ->
// subspace matrix -> the reconstruction loop -> error_history.json and one
// .cfl per volume named <prefix><volume index>, zero-padded to the digit count
// of the data set's volume count. The block shifts of the reference are
// unseeded random draws in [0, max_rand_shift) (lib.rs:420-427); here they come
// from a seeded generator (--seed, else the TOML key shift_seed, else 0) and
// are written to <output_cfl_dir>/shifts.txt so a run can be repeated.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/llr_hip.h"

namespace {

int die(const std::string& msg) {
  std::fprintf(stderr, "llr-recon: %s\n", msg.c_str());
  return 1;
}

bool check(llr_status st, const char* what) {
  if (st == LLR_OK) return true;
  std::fprintf(stderr, "llr-recon: %s failed (status %d): %s\n", what, (int)st, llr_last_error());
  return false;
}

void usage(FILE* f) {
  std::fprintf(f,
               "usage: llr-recon <recon_params> <input_cfl_dir> <input_cfl_pattern> <output_cfl_dir>\n"
               "                 <output_cfl_prefix> <b_table> [b0_null_vec_tol] [sample_mask_cfl]\n"
               "                 [expected_volumes] [is_image] [recon_b0] [--seed N] [--device N]\n");
}

bool parse_bool(const std::string& s, bool* out) {
  if (s == "true") {
    *out = true;
    return true;
  }
  if (s == "false") {
    *out = false;
    return true;
  }
  return false;
}

}  // namespace

int main(int argc, char** argv) {
  std::vector<std::string> pos;
  long long seed_arg = -1;
  int device = 0;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "-h" || a == "--help") {
      usage(stdout);
      return 0;
    }
    if (a == "--seed" || a == "--device") {
      if (i + 1 >= argc) return die(a + " needs a value");
      char* end = nullptr;
      const long long v = std::strtoll(argv[++i], &end, 10);
      if (*end != '\0' || v < 0) return die("bad value for " + a);
      if (a == "--seed")
        seed_arg = v;
      else
        device = (int)v;
      continue;
    }
    pos.push_back(a);
  }
  if (pos.size() < 6 || pos.size() > 11) {
    usage(stderr);
    return 2;
  }
  const std::string params_path = pos[0], in_dir = pos[1], pattern = pos[2], out_dir = pos[3], prefix = pos[4],
                    btable = pos[5];
  float b0_tol = 0.f;
  std::string mask_path;
  size_t expected = 0;
  bool is_image = false, recon_b0 = false;
  if (pos.size() > 6) {
    char* end = nullptr;
    b0_tol = std::strtof(pos[6].c_str(), &end);
    if (*end != '\0') return die("b0_null_vec_tol: not a number: " + pos[6]);
  }
  if (pos.size() > 7) mask_path = pos[7];
  if (pos.size() > 8) {
    char* end = nullptr;
    const long long e = std::strtoll(pos[8].c_str(), &end, 10);
    if (*end != '\0' || e < 0) return die("expected_volumes: not a count: " + pos[8]);
    expected = (size_t)e;
  }
  if (pos.size() > 9 && !parse_bool(pos[9], &is_image)) return die("is_image: expected true or false");
  if (pos.size() > 10 && !parse_bool(pos[10], &recon_b0)) return die("recon_b0: expected true or false");

  llr_toml_params rp;
  if (!check(llr_recon_params_read(params_path.c_str(), &rp), "reading the recon params")) return 1;

  llr_dataset* ds = nullptr;
  if (!check(llr_dataset_from_cfl_dir(in_dir.c_str(), pattern.c_str(), expected, &ds), "finding the volumes")) return 1;
  struct DsGuard {
    llr_dataset* d;
    ~DsGuard() { llr_dataset_destroy(d); }
  } dsg{ds};
  if (!check(llr_dataset_with_b_table(ds, btable.c_str(), b0_tol), "reading the b-table")) return 1;
  if (!mask_path.empty() && !check(llr_dataset_with_sample_mask(ds, mask_path.c_str()), "reading the sample mask"))
    return 1;
  size_t vol[3], n_total = 0;
  llr_dataset_info(ds, vol, &n_total, nullptr, nullptr);
  const int which = recon_b0 ? LLR_VOLUMES_B0 : LLR_VOLUMES_DTI;
  size_t nq = 0;
  llr_dataset_indices(ds, which, nullptr, 0, &nq);
  if (nq == 0) return die(recon_b0 ? "no b0 volumes in the b-table" : "no diffusion-weighted volumes in the b-table");
  std::vector<size_t> vidx(nq);
  llr_dataset_indices(ds, which, vidx.data(), nq, &nq);
  const size_t N = vol[0] * vol[1] * vol[2];
  std::printf("loading data ...\n");
  std::vector<llr_cf32> data(N * nq);
  if (!check(llr_dataset_load(ds, which, data.data(), data.size()), "loading the volumes")) return 1;

  // lib.rs:196-207: tall blocks, 0 < rank <= columns
  const size_t m = rp.llr_block_size[0] * rp.llr_block_size[1] * rp.llr_block_size[2];
  if (m <= nq) return die("only tall matrices are supported");

  llr_ctx* ctx = nullptr;
  if (!check(llr_ctx_create(device, &ctx), "creating the device context")) return 1;
  struct CtxGuard {
    llr_ctx* c;
    ~CtxGuard() { llr_ctx_destroy(c); }
  } cg{ctx};

  if (is_image) {  // lib.rs:151-168: image data through the forward model
    std::printf("converting image data to k-space ...\n");
    std::vector<float> zero(3 * nq, 0.f);
    if (!check(llr_signal_model(ctx, data.data(), data.size(), vol, nq, zero.data(), zero.data(), LLR_FORWARD),
               "the forward model"))
      return 1;
  }
  if (!mask_path.empty()) {  // lib.rs:171-179
    std::vector<llr_cf32> mask(vol[1] * vol[2] * nq);
    if (!check(llr_dataset_sample_mask(ds, which, mask.data(), mask.size()), "selecting the sample masks")) return 1;
    if (!check(llr_under_sample_kspace(ctx, data.data(), vol, nq, mask.data()), "undersampling")) return 1;
  }

  // subspace matrix (lib.rs:210-219): an nq x rank .cfl with 2 non-singleton dims
  std::vector<llr_cf32> sub;
  size_t srank = 0;
  if (rp.has_subspace) {
    std::printf("loading subspace matrix ...\n");
    size_t d[16], nd = 0;
    if (!check(llr_cfl_dims(rp.subspace_projection_mat, d, &nd), "reading the subspace matrix header")) return 1;
    std::vector<size_t> ns;
    for (size_t i = 0; i < nd; ++i)
      if (d[i] != 1) ns.push_back(d[i]);
    if (ns.size() != 2) return die("matrix to have 2 non-singleton dimensions");
    if (ns[0] != nq) return die("unexpected subspace dimension");
    srank = ns[1];
    sub.resize(nq * srank);
    if (!check(llr_cfl_read(rp.subspace_projection_mat, sub.data(), sub.size()), "reading the subspace matrix"))
      return 1;
    std::printf("loaded subspace matrix with rank %zu\n", srank);
    if (rp.llr_mode == LLR_MODE_RANK && srank < rp.llr_block_rank)
      return die("llr rank must be less than or equal to subspace rank");
  }
  if (rp.llr_mode == LLR_MODE_RANK) {
    if (rp.llr_block_rank > (srank ? srank : nq)) return die("rank must not exceed n");
    if (rp.llr_block_rank == 0) return die("rank should not be less than 1");
  }

  // block shifts (lib.rs:260, 420-427): uniform in [0, max_rand_shift) per axis
  size_t maxs[3];
  for (int a = 0; a < 3; ++a) maxs[a] = rp.has_max_rand_shift ? rp.max_rand_shift[a] : vol[a];
  for (int a = 0; a < 3; ++a)
    if (maxs[a] == 0) return die("max_rand_shift entries must be positive");
  const unsigned long long seed = seed_arg >= 0 ? (unsigned long long)seed_arg : rp.seed;
  std::mt19937_64 gen(seed);
  std::vector<size_t> shifts(3 * rp.n_iter);
  for (size_t it = 0; it < rp.n_iter; ++it)
    for (int a = 0; a < 3; ++a) shifts[3 * it + a] = std::uniform_int_distribution<size_t>(0, maxs[a] - 1)(gen);

  llr_recon_params p;
  std::memset(&p, 0, sizeof p);
  for (int a = 0; a < 3; ++a) {
    p.vol[a] = vol[a];
    p.blk[a] = rp.llr_block_size[a];
  }
  p.nq = nq;
  p.n_iter = rp.n_iter;
  p.llr_mode = rp.llr_mode;
  p.precision = rp.precision;
  p.rank = rp.llr_block_rank;
  p.lambda_factor = rp.lambda_factor;
  p.wavelet_alpha = rp.wavelet_shrink_factor;
  p.wavelet_levels = -1;
  p.srank = srank;
  p.subspace = srank ? sub.data() : nullptr;
  p.shifts = rp.n_iter ? shifts.data() : nullptr;
  p.storage = rp.storage;
  llr_recon* r = nullptr;
  if (!check(llr_recon_create(ctx, &p, &r), "creating the reconstruction")) return 1;
  struct RGuard {
    llr_recon* r;
    ~RGuard() { llr_recon_destroy(r); }
  } rg{r};
  if (!check(llr_recon_set_measurements(r, data.data(), data.size(), 0), "uploading the measurements")) return 1;
  std::printf("running %zu iterations on %zux%zux%zu x %zu volumes ...\n", rp.n_iter, vol[0], vol[1], vol[2], nq);
  std::vector<double> hist(rp.n_iter ? rp.n_iter : 1);
  size_t done = 0;
  if (!check(llr_recon_run(r, hist.data(), &done), "the reconstruction")) return 1;
  for (size_t i = 0; i < done; ++i) std::printf("iteration %zu: mse / total energy %.3e\n", i + 1, hist[i]);
  if (!check(llr_recon_get_image(r, data.data(), data.size(), 0), "downloading the image")) return 1;

  // outputs (lib.rs:345-362)
  std::string od = out_dir;
  if (!od.empty() && od.back() != '/') od += '/';
  if (!check(llr_write_error_history((od + "error_history.json").c_str(), hist.data(), done), "error history"))
    return 1;
  if (FILE* f = std::fopen((od + "shifts.txt").c_str(), "w")) {
    for (size_t it = 0; it < rp.n_iter; ++it)
      std::fprintf(f, "%zu %zu %zu\n", shifts[3 * it], shifts[3 * it + 1], shifts[3 * it + 2]);
    std::fclose(f);
  } else {
    return die("failed to write " + od + "shifts.txt");
  }
  std::printf("writing results ...\n");
  const int width = (int)std::to_string(n_total).size();
  for (size_t k = 0; k < nq; ++k) {
    char name[64];
    std::snprintf(name, sizeof name, "%0*zu", width, vidx[k]);
    if (!check(llr_cfl_write((od + prefix + name).c_str(), vol, 3, data.data() + k * N, N), "writing a volume"))
      return 1;
  }
  return 0;
}
