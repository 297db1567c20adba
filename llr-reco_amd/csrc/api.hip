// C ABI of the LLR HIP library: host-pointer drop-ins for the reference's Rust
// API and the device-resident reconstruction loop (llr_recon_exec,
// lib.rs:113-363). Declarations and reference citations are in llr_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "common.h"
#include "internal.h"

namespace llr {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

llr_status fail(llr_status st, const std::string& msg) {
  g_err = msg;
  return st;
}

llr_status hip_fail(hipError_t e, const char* what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return e == hipErrorOutOfMemory ? LLR_ERR_NOMEM : LLR_ERR_HIP;
}

llr_status get_twiddles(llr_ctx* ctx, int n, const float2** out) {
  for (auto& p : ctx->twiddles)
    if (p.first == n) {
      *out = p.second;
      return LLR_OK;
    }
  std::vector<float2> h(n);
  const long double TWO_PI = 6.283185307179586476925286766559005768L;
  for (int m = 0; m < n; ++m) {
    long double ang = -TWO_PI * (long double)m / (long double)n;
    h[m] = make_float2((float)std::cos(ang), (float)std::sin(ang));
  }
  float2* d = nullptr;
  LLR_HIP(hipMalloc(&d, sizeof(float2) * n));
  LLR_HIP(hipMemcpy(d, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
  ctx->twiddles.push_back({n, d});
  *out = d;
  return LLR_OK;
}

llr_status red_scratch(llr_ctx* ctx, size_t n, double** out) {
  if (ctx->red_cap < n) {
    if (ctx->d_red) (void)hipFree(ctx->d_red);
    ctx->d_red = nullptr;
    LLR_HIP(hipMalloc(&ctx->d_red, n * sizeof(double)));
    ctx->red_cap = n;
  }
  *out = ctx->d_red;
  return LLR_OK;
}

// exp(+i * 2*pi * frac((i - n/2) * s / n)), s = c + n/2 (signal_model.rs:110-117,
// cufft.rs:320-326), reduced to turns in long double.
void ramp_table(std::vector<float2>& out, size_t n, const float* corr, size_t stride, size_t nvols) {
  out.resize(n * nvols);
  const long double TWO_PI = 6.283185307179586476925286766559005768L;
  for (size_t v = 0; v < nvols; ++v) {
    long double s = (long double)(corr ? corr[v * stride] : 0.0f) + 0.5L * (long double)n;
    for (size_t i = 0; i < n; ++i) {
      long double tr = ((long double)i - 0.5L * (long double)n) * s / (long double)n;
      tr -= std::floor(tr);
      long double ang = TWO_PI * tr;
      out[v * n + i] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    }
  }
}

// ------------------------------------------------------------- helpers
static llr_status upload(DevBuf& b, const void* h, size_t bytes) {
  if (!(b.p && b.bytes == bytes)) LLR_HIP(b.alloc(bytes));  // reuse a buffer of the same size
  if (h && bytes) LLR_HIP(hipMemcpy(b.p, h, bytes, hipMemcpyHostToDevice));
  return LLR_OK;
}

static llr_status download(void* h, const DevBuf& b, size_t bytes, llr_ctx* ctx) {
  LLR_HIP(hipStreamSynchronize(ctx->stream));
  if (bytes) LLR_HIP(hipMemcpy(h, b.p, bytes, hipMemcpyDeviceToHost));
  return LLR_OK;
}

struct RampTables {
  DevBuf pk[3], pi[3];
  // pi[a] is the parity pattern (-1)^(i + n/2): axis length n even and no
  // image correction on that axis (ramp_table's angle is then 0 or pi; the
  // imaginary parts it stores are below 1e-19)
  bool pi_sign[3] = {false, false, false};
};

static llr_status make_tables(RampTables& t, const size_t* dims, size_t nv, const float* ksp, const float* img) {
  for (int a = 0; a < 3; ++a) {
    t.pi_sign[a] = dims[a] % 2 == 0;
    if (img)
      for (size_t v = 0; v < nv; ++v) t.pi_sign[a] = t.pi_sign[a] && img[3 * v + a] == 0.f;
    std::vector<float2> h;
    ramp_table(h, dims[a], ksp ? ksp + a : nullptr, 3, nv);
    llr_status st = upload(t.pk[a], h.data(), h.size() * sizeof(float2));
    if (st) return st;
    ramp_table(h, dims[a], img ? img + a : nullptr, 3, nv);
    st = upload(t.pi[a], h.data(), h.size() * sizeof(float2));
    if (st) return st;
  }
  return LLR_OK;
}

static llr_status check_fft_dims(const size_t* dims) {
  for (int a = 0; a < 3; ++a)
    if (dims[a] < 2 || !fft_size_supported(dims[a]))
      return fail(LLR_ERR_UNSUPPORTED, "volume dimension " + std::to_string(dims[a]) +
                                           " is not a supported FFT length (2..512, radix 2/3)");
  return LLR_OK;
}

// signal model (signal_model.rs:34-79) on device data; src -> dst (may alias)
// (src and dst both complex64, or both half-precision pairs when h16 is set)
// (axes a0..2: a0 = 1 leaves the x axis alone)
static llr_status model_device(llr_ctx* ctx, const void* src, void* dst, const size_t* dims, size_t nv,
                               const RampTables& t, int dir, bool h16 = false, int a0 = 0) {
  const void* in = src;
  for (int a = a0; a < 3; ++a) {
    float s = (float)(1.0 / std::sqrt((double)dims[a]));
    llr_status st;
    if (dir == LLR_FORWARD)
      st = fft_axis_pass(ctx, in, dst, dims, nv, a, false, t.pk[a].as<float2>(), false, t.pi[a].as<float2>(), true, s,
                         h16);
    else
      st = fft_axis_pass(ctx, in, dst, dims, nv, a, true, t.pi[a].as<float2>(), false, t.pk[a].as<float2>(), true, s,
                         h16);
    if (st) return st;
    in = dst;
  }
  return LLR_OK;
}

static llr_status check_blocks(const size_t* vol, const size_t* blk) {
  for (int a = 0; a < 3; ++a) {
    if (blk[a] == 0) return fail(LLR_ERR_ARG, "block dimension must be positive");
    if (blk[a] > vol[a]) return fail(LLR_ERR_ARG, "block dimension must be less than or equal to volume dimension");
  }
  return LLR_OK;
}

}  // namespace llr

using namespace llr;

// ===================================================================== ABI
extern "C" {

const char* llr_last_error(void) { return g_err.c_str(); }
int llr_abi_version(void) { return 108; }

llr_status llr_ctx_create(int device, llr_ctx** out) {
  if (!out) return fail(LLR_ERR_ARG, "null output pointer");
  int n = 0;
  LLR_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(LLR_ERR_ARG, "no such HIP device: " + std::to_string(device));
  LLR_HIP(hipSetDevice(device));
  llr_ctx* c = new llr_ctx();
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return hip_fail(e, "hipStreamCreate");
  }
  *out = c;
  return LLR_OK;
}

void llr_ctx_destroy(llr_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& p : ctx->twiddles) (void)hipFree(p.second);
  if (ctx->d_red) (void)hipFree(ctx->d_red);
  for (hipEvent_t e : ctx->kt_ev)
    if (e) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

llr_status llr_ctx_synchronize(llr_ctx* ctx) {
  LLR_HIP(hipStreamSynchronize(ctx->stream));
  return LLR_OK;
}

int llr_ctx_device(const llr_ctx* ctx) { return ctx ? ctx->device : -1; }

void llr_grid_dim(const size_t vol_size[3], const size_t block_size[3], size_t grid[3]) {
  for (int a = 0; a < 3; ++a) grid[a] = block_size[a] ? vol_size[a] / block_size[a] : 0;
}

// ------------------------------------------------------------------ block.rs
static llr_status lift_common(llr_ctx* ctx, const void* volumes, void* vol_out, size_t nvols, size_t eb,
                              const void* mats_in, void* mats_out, size_t matrix_len, const size_t* vol,
                              const size_t* blk, const size_t* shift, size_t start, size_t n, bool unlift) {
  if (!ctx) return fail(LLR_ERR_ARG, "null context");
  if (eb == 0) return fail(LLR_ERR_ARG, "element size must be positive");
  llr_status st = check_blocks(vol, blk);
  if (st) return st;
  size_t m = blk[0] * blk[1] * blk[2];
  size_t need = m * nvols * n;
  if (matrix_len < need) return fail(LLR_ERR_ARG, "matrix data buffer is too small");
  size_t grid[3];
  llr_grid_dim(vol, blk, grid);
  if (start + n > grid[0] * grid[1] * grid[2]) return fail(LLR_ERR_ARG, "matrix range past the block grid");
  size_t N = vol[0] * vol[1] * vol[2];
  LLR_HIP(hipSetDevice(ctx->device));
  DevBuf dv, dm;
  st = upload(dv, volumes, N * nvols * eb);
  if (st) return st;
  st = upload(dm, unlift ? mats_in : nullptr, need * eb);
  if (st) return st;
  st = lift_device(ctx, dv.p, dm.p, nvols, eb, vol, blk, shift, start, n, unlift);
  if (st) return st;
  if (unlift) return download(vol_out, dv, N * nvols * eb, ctx);
  return download(mats_out, dm, need * eb, ctx);
}

llr_status llr_lift_data_set(llr_ctx* ctx, const void* volumes, size_t n_volumes, size_t elem_bytes, void* matrix_data,
                             size_t matrix_len, const size_t vol_size[3], const size_t block_size[3],
                             const size_t shift[3], size_t matrix_idx_start, size_t n_matrices) {
  return lift_common(ctx, volumes, nullptr, n_volumes, elem_bytes, nullptr, matrix_data, matrix_len, vol_size,
                     block_size, shift, matrix_idx_start, n_matrices, false);
}

llr_status llr_un_lift_data_set(llr_ctx* ctx, void* volumes, size_t n_volumes, size_t elem_bytes,
                                const void* matrix_data, size_t matrix_len, const size_t vol_size[3],
                                const size_t block_size[3], const size_t shift[3], size_t matrix_idx_start,
                                size_t n_matrices) {
  return lift_common(ctx, volumes, volumes, n_volumes, elem_bytes, matrix_data, nullptr, matrix_len, vol_size,
                     block_size, shift, matrix_idx_start, n_matrices, true);
}

static llr_status lift_device_common(llr_ctx* ctx, const void* volumes, size_t nvols, size_t eb, void* mats,
                                     size_t matrix_len, const size_t* vol, const size_t* blk, const size_t* shift,
                                     size_t start, size_t n, bool unlift) {
  if (!ctx) return fail(LLR_ERR_ARG, "null context");
  if (eb == 0) return fail(LLR_ERR_ARG, "element size must be positive");
  if (!volumes || !mats) return fail(LLR_ERR_ARG, "null buffer");
  llr_status st = check_blocks(vol, blk);
  if (st) return st;
  const size_t m = blk[0] * blk[1] * blk[2];
  if (matrix_len < m * nvols * n) return fail(LLR_ERR_ARG, "matrix data buffer is too small");
  size_t grid[3];
  llr_grid_dim(vol, blk, grid);
  if (start + n > grid[0] * grid[1] * grid[2]) return fail(LLR_ERR_ARG, "matrix range past the block grid");
  LLR_HIP(hipSetDevice(ctx->device));
  if ((st = lift_device(ctx, volumes, mats, nvols, eb, vol, blk, shift, start, n, unlift))) return st;
  LLR_HIP(hipStreamSynchronize(ctx->stream));
  return LLR_OK;
}

llr_status llr_lift_data_set_device(llr_ctx* ctx, const void* volumes, size_t n_volumes, size_t elem_bytes,
                                    void* matrix_data, size_t matrix_len, const size_t vol_size[3],
                                    const size_t block_size[3], const size_t shift[3], size_t matrix_idx_start,
                                    size_t n_matrices) {
  return lift_device_common(ctx, volumes, n_volumes, elem_bytes, matrix_data, matrix_len, vol_size, block_size, shift,
                            matrix_idx_start, n_matrices, false);
}

llr_status llr_un_lift_data_set_device(llr_ctx* ctx, void* volumes, size_t n_volumes, size_t elem_bytes,
                                       const void* matrix_data, size_t matrix_len, const size_t vol_size[3],
                                       const size_t block_size[3], const size_t shift[3], size_t matrix_idx_start,
                                       size_t n_matrices) {
  return lift_device_common(ctx, volumes, n_volumes, elem_bytes, (void*)matrix_data, matrix_len, vol_size,
                            block_size, shift, matrix_idx_start, n_matrices, true);
}

// ----------------------------------------------------------- dti_subspace.rs
llr_status llr_dti_project(llr_ctx* ctx, const llr_cf32* subspace, size_t subspace_len, const llr_cf32* image_matrices,
                           size_t image_len, llr_cf32* compressed, size_t compressed_len, size_t batch_size, size_t nq,
                           size_t nvox, size_t srank) {
  if (image_len < nq * nvox * batch_size) return fail(LLR_ERR_ARG, "image matrices doesn't have enough entries");
  if (subspace_len != nq * srank) return fail(LLR_ERR_ARG, "unexpected subspace matrix size");
  if (compressed_len != nvox * srank * batch_size)
    return fail(LLR_ERR_ARG, "compressed result doesn't have enough entries");
  LLR_HIP(hipSetDevice(ctx->device));
  DevBuf dS, dX, dC;
  llr_status st;
  if ((st = upload(dS, subspace, subspace_len * 8))) return st;
  if ((st = upload(dX, image_matrices, nq * nvox * batch_size * 8))) return st;
  if ((st = upload(dC, nullptr, compressed_len * 8))) return st;
  st = cgemm_device(ctx, srank, nq, nvox, batch_size, true, make_float2(1, 0), make_float2(0, 0), true, true,
                    dS.as<float2>(), dX.as<float2>(), dC.as<float2>());
  if (st) return st;
  return download(compressed, dC, compressed_len * 8, ctx);
}

llr_status llr_dti_back_project(llr_ctx* ctx, const llr_cf32* subspace, size_t subspace_len, llr_cf32* image_matrices,
                                size_t image_len, const llr_cf32* compressed, size_t compressed_len, size_t batch_size,
                                size_t nq, size_t nvox, size_t srank) {
  if (image_len != nq * nvox * batch_size) return fail(LLR_ERR_ARG, "image matrices doesn't have enough entries");
  if (subspace_len != nq * srank) return fail(LLR_ERR_ARG, "unexpected subspace matrix size");
  if (compressed_len != nvox * srank * batch_size)
    return fail(LLR_ERR_ARG, "compressed result doesn't have enough entries");
  LLR_HIP(hipSetDevice(ctx->device));
  DevBuf dS, dB, dC, dX;
  llr_status st;
  if ((st = upload(dS, subspace, subspace_len * 8))) return st;
  if ((st = upload(dC, compressed, compressed_len * 8))) return st;
  if ((st = upload(dB, nullptr, image_len * 8))) return st;
  if ((st = upload(dX, nullptr, image_len * 8))) return st;
  st = cgemm_device(ctx, nq, srank, nvox, batch_size, true, make_float2(1, 0), make_float2(0, 0), false, false,
                    dS.as<float2>(), dC.as<float2>(), dB.as<float2>());
  if (st) return st;
  st = conj_transpose_device(ctx, dB.as<float2>(), dX.as<float2>(), nq, nvox, batch_size, true);
  if (st) return st;
  return download(image_matrices, dX, image_len * 8, ctx);
}

llr_status llr_conj_transpose_matrix(llr_ctx* ctx, llr_cf32* matrix_data, size_t len, size_t m, size_t n,
                                     size_t batch_size) {
  if (len != m * n * batch_size) return fail(LLR_ERR_ARG, "matrix data doesn't have enough entries");
  LLR_HIP(hipSetDevice(ctx->device));
  DevBuf a, b;
  llr_status st;
  if ((st = upload(a, matrix_data, len * 8))) return st;
  if ((st = upload(b, nullptr, len * 8))) return st;
  if ((st = conj_transpose_device(ctx, a.as<float2>(), b.as<float2>(), m, n, batch_size, true))) return st;
  return download(matrix_data, b, len * 8, ctx);
}

llr_status llr_subspace_constraint(llr_ctx* ctx, const llr_cf32* subspace, size_t subspace_len, llr_cf32* image_matrices,
                                   size_t image_len, size_t batch_size, size_t nq, size_t nvox, size_t srank) {
  if (image_len < nq * nvox * batch_size) return fail(LLR_ERR_ARG, "image matrices doesn't have enough entries");
  if (subspace_len != nq * srank) return fail(LLR_ERR_ARG, "unexpected subspace matrix size");
  size_t xl = nq * nvox * batch_size;
  LLR_HIP(hipSetDevice(ctx->device));
  DevBuf dS, dX, dC, dB;
  llr_status st;
  if ((st = upload(dS, subspace, subspace_len * 8))) return st;
  if ((st = upload(dX, image_matrices, xl * 8))) return st;
  if ((st = upload(dC, nullptr, nvox * srank * batch_size * 8))) return st;
  if ((st = upload(dB, nullptr, xl * 8))) return st;
  if ((st = cgemm_device(ctx, srank, nq, nvox, batch_size, true, make_float2(1, 0), make_float2(0, 0), true, true,
                         dS.as<float2>(), dX.as<float2>(), dC.as<float2>())))
    return st;
  if ((st = cgemm_device(ctx, nq, srank, nvox, batch_size, true, make_float2(1, 0), make_float2(0, 0), false, false,
                         dS.as<float2>(), dC.as<float2>(), dB.as<float2>())))
    return st;
  // dti_subspace.rs:127-133 writes the plain transpose of (nq x nvox) into (nvox x nq)
  if ((st = conj_transpose_device(ctx, dB.as<float2>(), dX.as<float2>(), nq, nvox, batch_size, false))) return st;
  return download(image_matrices, dX, xl * 8, ctx);
}

// ------------------------------------------------- low_rank.rs / cublas.rs
static llr_status lowrank_common(llr_ctx* ctx, size_t m, size_t n, int mode, size_t rank, float lambda, size_t batch,
                                 llr_cf32* data, size_t len, float* smax_out) {
  if (len < m * n * batch) return fail(LLR_ERR_ARG, "matrix data buffer is too small");
  if (n == 0 || m < n) return fail(LLR_ERR_ARG, "only tall matrices (m >= n > 0) are supported");
  if (mode == 0 && (rank == 0 || rank > n)) return fail(LLR_ERR_ARG, "rank must be in 1..=n");
  if (batch == 0) return LLR_OK;
  LLR_HIP(hipSetDevice(ctx->device));
  size_t el = m * n * batch;
  DevBuf d, smax;
  llr_status st;
  if ((st = upload(d, data, el * 8))) return st;
  if ((st = upload(smax, nullptr, sizeof(unsigned)))) return st;
  LLR_HIP(hipMemsetAsync(smax.p, 0, sizeof(unsigned), ctx->stream));
  if (n <= 8) {
    size_t vol[3] = {m, 1, 1}, zero[3] = {0, 0, 0};
    st = llr_block_run(ctx, d.as<float2>(), d.as<float2>(), nullptr, 0, vol, vol, zero, (int)n, 1, 0, (long long)batch,
                       mode, (int)rank, lambda, smax.as<unsigned>(), nullptr);
  } else {
    st = lowrank_generic_device(ctx, d.as<float2>(), (long long)m, (int)n, (long long)batch, mode, (int)rank, lambda,
                                smax.as<unsigned>());
  }
  if (st) return st;
  if (smax_out) {
    unsigned bits = 0;
    LLR_HIP(hipStreamSynchronize(ctx->stream));
    LLR_HIP(hipMemcpy(&bits, smax.p, sizeof(unsigned), hipMemcpyDeviceToHost));
    std::memcpy(smax_out, &bits, sizeof(float));
  }
  return download(data, d, el * 8, ctx);
}

llr_status llr_low_rank_approx_batch(llr_ctx* ctx, size_t m, size_t n, size_t rank, size_t batch_size,
                                     llr_cf32* matrix_data, size_t len) {
  return lowrank_common(ctx, m, n, 0, rank, 0.f, batch_size, matrix_data, len, nullptr);
}

llr_status llr_svt_batch(llr_ctx* ctx, size_t m, size_t n, float lambda, size_t batch_size, llr_cf32* matrix_data,
                         size_t len, float* sigma_max) {
  if (!(lambda >= 0.f)) return fail(LLR_ERR_ARG, "lambda must be non-negative");
  return lowrank_common(ctx, m, n, 1, 0, lambda, batch_size, matrix_data, len, sigma_max);
}

llr_status llr_cgemm_strided_batched(llr_ctx* ctx, size_t m, size_t k, size_t n, size_t batch_size, int broadcast_a,
                                     llr_cf32 alpha, llr_cf32 beta, int transpose_a, int transpose_b, const llr_cf32* a,
                                     size_t a_len, const llr_cf32* b, size_t b_len, llr_cf32* c, size_t c_len) {
  size_t na = broadcast_a ? m * k : m * k * batch_size;
  if (a_len < na || b_len < k * n * batch_size || c_len < m * n * batch_size)
    return fail(LLR_ERR_ARG, "matrix buffers are too small");
  LLR_HIP(hipSetDevice(ctx->device));
  DevBuf dA, dB, dC;
  llr_status st;
  if ((st = upload(dA, a, na * 8))) return st;
  if ((st = upload(dB, b, k * n * batch_size * 8))) return st;
  if ((st = upload(dC, c, m * n * batch_size * 8))) return st;
  if ((st = cgemm_device(ctx, m, k, n, batch_size, broadcast_a != 0, make_float2(alpha.re, alpha.im),
                         make_float2(beta.re, beta.im), transpose_a != 0, transpose_b != 0, dA.as<float2>(),
                         dB.as<float2>(), dC.as<float2>())))
    return st;
  return download(c, dC, m * n * batch_size * 8, ctx);
}

llr_status llr_rhs_cdgmm_batched(llr_ctx* ctx, size_t m, size_t n, size_t batch_size, llr_cf32* matrix, size_t len,
                                 const float* diag) {
  if (len < m * n * batch_size) return fail(LLR_ERR_ARG, "matrix buffer is too small");
  LLR_HIP(hipSetDevice(ctx->device));
  DevBuf dA, dD;
  llr_status st;
  if ((st = upload(dA, matrix, m * n * batch_size * 8))) return st;
  if ((st = upload(dD, diag, n * batch_size * 4))) return st;
  if ((st = cdgmm_device(ctx, dA.as<float2>(), dD.as<float>(), m, n, batch_size))) return st;
  return download(matrix, dA, m * n * batch_size * 8, ctx);
}

// ----------------------------------------------- signal_model.rs / cufft.rs
llr_status llr_signal_model(llr_ctx* ctx, llr_cf32* x, size_t len, const size_t image_size[3], size_t batch_size,
                            const float* lin_ksp_corr, const float* lin_img_corr, int direction) {
  size_t N = image_size[0] * image_size[1] * image_size[2];
  if (len != N * batch_size) return fail(LLR_ERR_ARG, "data layout and data length inconsistency");
  if (direction != LLR_FORWARD && direction != LLR_INVERSE) return fail(LLR_ERR_ARG, "bad direction");
  llr_status st = check_fft_dims(image_size);
  if (st) return st;
  LLR_HIP(hipSetDevice(ctx->device));
  RampTables t;
  if ((st = make_tables(t, image_size, batch_size, lin_ksp_corr, lin_img_corr))) return st;
  DevBuf d;
  if ((st = upload(d, x, len * 8))) return st;
  if ((st = model_device(ctx, d.as<float2>(), d.as<float2>(), image_size, batch_size, t, direction))) return st;
  return download(x, d, len * 8, ctx);
}

llr_status llr_signal_model_device(llr_ctx* ctx, llr_cf32* x, size_t len, const size_t image_size[3],
                                   size_t batch_size, const float* lin_ksp_corr, const float* lin_img_corr,
                                   int direction) {
  size_t N = image_size[0] * image_size[1] * image_size[2];
  if (len != N * batch_size) return fail(LLR_ERR_ARG, "data layout and data length inconsistency");
  if (direction != LLR_FORWARD && direction != LLR_INVERSE) return fail(LLR_ERR_ARG, "bad direction");
  llr_status st = check_fft_dims(image_size);
  if (st) return st;
  LLR_HIP(hipSetDevice(ctx->device));
  RampTables t;
  if ((st = make_tables(t, image_size, batch_size, lin_ksp_corr, lin_img_corr))) return st;
  if ((st = model_device(ctx, (float2*)x, (float2*)x, image_size, batch_size, t, direction))) return st;
  LLR_HIP(hipStreamSynchronize(ctx->stream));
  return LLR_OK;
}

llr_status llr_estimate_linear_correction(llr_ctx* ctx, const llr_cf32* ksp, size_t len, const size_t vol_size[3],
                                          float out[3]) {
  size_t N = vol_size[0] * vol_size[1] * vol_size[2];
  if (len != N) return fail(LLR_ERR_ARG, "mismatch between vol size and number of samples");
  LLR_HIP(hipSetDevice(ctx->device));
  DevBuf d, best;
  llr_status st;
  if ((st = upload(d, ksp, len * 8))) return st;
  if ((st = upload(best, nullptr, 8))) return st;
  LLR_HIP(hipMemsetAsync(best.p, 0, 8, ctx->stream));
  if ((st = argmax_volumes(ctx, d.as<float2>(), N, 1, best.as<unsigned long long>()))) return st;
  unsigned long long key = 0;
  if ((st = download(&key, best, 8, ctx))) return st;
  size_t idx = (size_t)(0xFFFFFFFFu - (unsigned)(key & 0xFFFFFFFFull));
  size_t s[3] = {idx % vol_size[0], (idx / vol_size[0]) % vol_size[1], idx / (vol_size[0] * vol_size[1])};
  for (int a = 0; a < 3; ++a) out[a] = (float)((long long)s[a] - ((long long)vol_size[a] + 1) / 2);
  return LLR_OK;
}

llr_status llr_fftn_batch(llr_ctx* ctx, const size_t dims[3], size_t batches, int direction, int norm, llr_cf32* data,
                          size_t len) {
  size_t N = dims[0] * dims[1] * dims[2];
  if (len != N * batches) return fail(LLR_ERR_ARG, "data layout and data length inconsistency");
  llr_status st = check_fft_dims(dims);
  if (st) return st;
  LLR_HIP(hipSetDevice(ctx->device));
  DevBuf d;
  if ((st = upload(d, data, len * 8))) return st;
  for (int a = 0; a < 3; ++a) {
    double s = 1.0;
    if (norm == LLR_NORM_UNITARY) s = 1.0 / std::sqrt((double)dims[a]);
    if (norm == LLR_NORM_INVERSE && direction == LLR_INVERSE) s = 1.0 / (double)dims[a];
    if ((st = fft_axis_pass(ctx, d.as<float2>(), d.as<float2>(), dims, batches, a, direction == LLR_INVERSE, nullptr,
                            false, nullptr, false, (float)s)))
      return st;
  }
  return download(data, d, len * 8, ctx);
}

// Centered unitary FFT (cufft.rs:340-402): phase_shift3(Inverse) -> unitary FFT ->
// phase_shift3(Forward). The modulation pi (i - n/2) per axis is the signal
// model's ramp with zero corrections (s = n/2), so this is the signal model's
// pass sequence with zero-correction tables, a chunk of volumes at a time.
llr_status llr_fft3c(llr_ctx* ctx, llr_cf32* x, size_t len, const size_t dims[3], size_t batch_size, int direction,
                     size_t vols_per_batch) {
  if (!ctx || !x || !dims) return fail(LLR_ERR_ARG, "null argument");
  const size_t N = dims[0] * dims[1] * dims[2];
  if (len != N * batch_size) return fail(LLR_ERR_ARG, "data layout and data length inconsistency");
  if (direction != LLR_FORWARD && direction != LLR_INVERSE) return fail(LLR_ERR_ARG, "bad direction");
  llr_status st = check_fft_dims(dims);
  if (st) return st;
  if (batch_size == 0) return LLR_OK;
  const size_t chunk = vols_per_batch ? std::min(vols_per_batch, batch_size) : batch_size;
  LLR_HIP(hipSetDevice(ctx->device));
  RampTables t;
  if ((st = make_tables(t, dims, chunk, nullptr, nullptr))) return st;
  DevBuf d;
  for (size_t v0 = 0; v0 < batch_size; v0 += chunk) {
    const size_t nv = std::min(chunk, batch_size - v0);
    if ((st = upload(d, x + v0 * N, nv * N * 8))) return st;
    if ((st = model_device(ctx, d.p, d.p, dims, nv, t, direction))) return st;
    if ((st = download(x + v0 * N, d, nv * N * 8, ctx))) return st;
  }
  return LLR_OK;
}

llr_status llr_cu_fft3(llr_ctx* ctx, llr_cf32* x, size_t len, const size_t dims[3], int direction) {
  if (direction != LLR_FORWARD && direction != LLR_INVERSE) return fail(LLR_ERR_ARG, "bad direction");
  return llr_fftn_batch(ctx, dims, 1, direction, LLR_NORM_INVERSE, x, len);
}

// ---------------------------------------------------------------- fftshift.rs
static llr_status shift_common(llr_ctx* ctx, void* data, size_t eb, const size_t* dims, bool inverse) {
  size_t N = dims[0] * dims[1] * dims[2];
  if (eb == 0 || N == 0) return fail(LLR_ERR_ARG, "empty data");
  long long sh[3];
  for (int a = 0; a < 3; ++a) sh[a] = inverse ? (long long)((dims[a] + 1) / 2) : (long long)(dims[a] / 2);
  LLR_HIP(hipSetDevice(ctx->device));
  DevBuf a, b;
  llr_status st;
  if ((st = upload(a, data, N * eb))) return st;
  if ((st = upload(b, nullptr, N * eb))) return st;
  if ((st = circshift_device(ctx, a.p, b.p, eb, dims, sh))) return st;
  return download(data, b, N * eb, ctx);
}

llr_status llr_fftshift(llr_ctx* ctx, void* data, size_t elem_bytes, const size_t dims[3]) {
  return shift_common(ctx, data, elem_bytes, dims, false);
}
llr_status llr_ifftshift(llr_ctx* ctx, void* data, size_t elem_bytes, const size_t dims[3]) {
  return shift_common(ctx, data, elem_bytes, dims, true);
}

llr_status llr_phase_shift3(llr_ctx* ctx, llr_cf32* x, size_t len, const size_t dims[3], int direction) {
  if (len != dims[0] * dims[1] * dims[2]) return fail(LLR_ERR_ARG, "dims must agree with the size of x");
  LLR_HIP(hipSetDevice(ctx->device));
  DevBuf d;
  llr_status st;
  if ((st = upload(d, x, len * 8))) return st;
  if ((st = phase_shift3_device(ctx, d.as<float2>(), dims, direction == LLR_FORWARD ? -1 : 1))) return st;
  return download(x, d, len * 8, ctx);
}

// --------------------------------------------------------------- lib.rs helpers
llr_status llr_wavelet_denoise(llr_ctx* ctx, llr_cf32* volumes, size_t n_volumes, const size_t vol_size[3], float alpha,
                               int levels) {
  size_t N = vol_size[0] * vol_size[1] * vol_size[2];
  if (N == 0 || n_volumes == 0) return LLR_OK;
  int L = levels < 0 ? wavelet_max_level(vol_size) : levels;
  size_t ce = 0, te = 0;
  wavelet_workspace_elems(vol_size, L, &ce, &te);
  LLR_HIP(hipSetDevice(ctx->device));
  DevBuf d, coef, t1, t2;
  llr_status st;
  if ((st = upload(d, volumes, N * n_volumes * 8))) return st;
  if ((st = upload(coef, nullptr, ce * n_volumes * 8))) return st;
  if ((st = upload(t1, nullptr, te * n_volumes * 8))) return st;
  if ((st = upload(t2, nullptr, te * n_volumes * 8))) return st;
  if ((st = wavelet_denoise_device(ctx, d.as<float2>(), n_volumes, vol_size, alpha, L, coef.as<float2>(),
                                   t1.as<float2>(), t2.as<float2>())))
    return st;
  return download(volumes, d, N * n_volumes * 8, ctx);
}

llr_status llr_hard_project(llr_ctx* ctx, llr_cf32* x, const llr_cf32* y, size_t len, double* sum_sq_error) {
  LLR_HIP(hipSetDevice(ctx->device));
  DevBuf dx, dy, part, out;
  const size_t np = 1024;
  llr_status st;
  if ((st = upload(dx, x, len * 8))) return st;
  if ((st = upload(dy, y, len * 8))) return st;
  if ((st = upload(part, nullptr, np * 8))) return st;
  if ((st = upload(out, nullptr, 8))) return st;
  if ((st = hard_project_partials(ctx, dx.as<float2>(), dy.as<float2>(), len, part.as<double>(), np))) return st;
  if ((st = reduce_strided(ctx, part.as<double>(), np, 1, out.as<double>()))) return st;
  double e = 0;
  if ((st = download(&e, out, 8, ctx))) return st;
  if (sum_sq_error) *sum_sq_error = e;
  return download(x, dx, len * 8, ctx);
}

llr_status llr_dti_scalars(llr_ctx* ctx, const llr_cf32* volumes, size_t n_voxels, size_t n_volumes, const float* b_vals,
                           const float* b_vecs, float b0_tol, float* scalars) {
  LLR_CHECK(ctx && volumes && b_vals && b_vecs && scalars, "null argument");
  LLR_HIP(hipSetDevice(ctx->device));
  DevBuf dv, dout;
  llr_status st;
  if ((st = upload(dv, volumes, n_voxels * n_volumes * 8))) return st;
  if ((st = upload(dout, nullptr, n_voxels * 6 * sizeof(float)))) return st;
  if ((st = dti_scalars_device(ctx, dv.as<float2>(), n_voxels, n_volumes, b_vals, b_vecs, b0_tol, dout.as<float>())))
    return st;
  return download(scalars, dout, n_voxels * 6 * sizeof(float), ctx);
}

llr_status llr_dti_scalars_device(llr_ctx* ctx, const llr_cf32* d_volumes, size_t n_voxels, size_t n_volumes,
                                  const float* b_vals, const float* b_vecs, float b0_tol, float* d_scalars) {
  LLR_CHECK(ctx && d_volumes && b_vals && b_vecs && d_scalars, "null argument");
  LLR_HIP(hipSetDevice(ctx->device));
  return dti_scalars_device(ctx, (const float2*)d_volumes, n_voxels, n_volumes, b_vals, b_vecs, b0_tol, d_scalars);
}

llr_status llr_calc_undersampling_factor(llr_ctx* ctx, const llr_cf32* y, size_t len, double* factor,
                                         double* total_energy) {
  LLR_HIP(hipSetDevice(ctx->device));
  DevBuf dy, part, out;
  const size_t np = 1024;
  llr_status st;
  if ((st = upload(dy, y, len * 8))) return st;
  if ((st = upload(part, nullptr, 2 * np * 8))) return st;
  if ((st = upload(out, nullptr, 16))) return st;
  if ((st = energy_partials(ctx, dy.as<float2>(), len, part.as<double>(), np))) return st;
  if ((st = reduce_strided(ctx, part.as<double>(), np, 2, out.as<double>()))) return st;
  double r[2];
  if ((st = download(r, out, 16, ctx))) return st;
  if (factor) *factor = (double)len / r[1];
  if (total_energy) *total_energy = r[0];
  return LLR_OK;
}

llr_status llr_under_sample_kspace(llr_ctx* ctx, llr_cf32* data, const size_t dims[3], size_t nq, const llr_cf32* mask) {
  size_t n = dims[0] * dims[1] * dims[2] * nq;
  LLR_HIP(hipSetDevice(ctx->device));
  DevBuf d, m;
  llr_status st;
  if ((st = upload(d, data, n * 8))) return st;
  if ((st = upload(m, mask, dims[1] * dims[2] * nq * 8))) return st;
  if ((st = under_sample_device(ctx, d.as<float2>(), dims, nq, m.as<float2>()))) return st;
  return download(data, d, n * 8, ctx);
}

}  // extern "C"

// ======================================================= device-resident recon
struct llr_recon {
  llr_ctx* ctx = nullptr;
  llr_recon_params p{};
  std::vector<size_t> shifts;
  size_t N = 0, nq = 0, D = 0;
  int nc = 0;       // columns of the LLR matrices (srank or nq)
  int wL = 0;       // wavelet levels
  DevBuf X, Y, Cb, S, coef, t1, t2, partials, stats, hist, smax, best, counters, werr;
  DevBuf Cprev;            // previous iteration's subspace coefficients (momentum)
  double momentum = 0.0;   // > 0: the LLR step acts on C(x_k) + momentum (C(x_k) - C(x_{k-1}))
  bool s_real = false;     // every subspace entry has a zero imaginary part
  DevBuf wide;             // more than 8 columns: per-block Gram / W of the wide LLR kernels
  bool mom_first = false;  // the next LLR step has no previous coefficients yet
  // half storage (LLR_STORE_FP16): X, coef hold half-precision pairs; Yh is the
  // copy of the measurements made by the set-up; Xin holds an initial image
  // (complex64) until the set-up converts it. Stored data are the caller's
  // times hs = sqrt(N / energy): no value can exceed sqrt(N) <= 2^14 (one value
  // is at most the norm of its volume, and every transform here preserves the
  // norm), and a typical image value is O(1).
  bool s16 = false, xin_pending = false;
  double hs = 1.0;
  // the block grid covers every voxel and the wavelet is on: level 0 of the
  // wavelet reads the LLR step's coefficient planes (coefK: their xy-bands)
  // and the back-projection is folded into it
  bool fuse_wavelet = false;
  DevBuf coefK;
  // Cb holds C = X S of the current iterate as project_current computes it
  // (left there by the set-up's threshold step)
  bool cb_current = false;
  DevBuf Yh, Xin;
  // hybrid: the measurements are a mask constant along x (every x-line is
  // fully measured or empty, before and after the transform), so the x-axis
  // transform commutes with the hard projection and the iteration works in
  // (x, ky, kz): the measurements are taken through the inverse x pass once
  // (Yx, complex64, or Yh in half storage) and the x passes are skipped
  bool hybrid = false;
  DevBuf Yx, flags;
  // fused adjoint projection (hybrid iteration with the fused subspace wavelet):
  // the adjoint's inverse y pass writes C = X S into Cb instead of X (cb_adj),
  // leaving X holding that pass's input (x_pending) until the next wavelet
  // rewrites X from the new coefficients, or the run ends and one inverse y
  // pass forms the image
  bool yproj = false, cb_adj = false, x_pending = false;
  RampTables tabs;
  unsigned long long llr_blocks = 0;  // blocks processed by the LLR step (mixed-form statistics)
  size_t parts_cap = 0;  // doubles in partials
  size_t vgroup = 0;     // volumes per group in the data-consistency passes (0: all at once)
  bool have_meas = false, have_init = false;
  bool prepared = false;  // set-up done (llr_recon_prepare / warm start) and not yet consumed by a run
  double lambda = 0.0, energy = 0.0;
  std::vector<float> ksp_corr;
  bool timing = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double stage_ms[LLR_NUM_STAGES] = {0};
  size_t stage_n[LLR_NUM_STAGES] = {0};
};

namespace {

struct StageTimer {
  llr_recon* r;
  int stage;
  bool count;  // false: add the time to the stage's current launch (a volume group after the first)
  StageTimer(llr_recon* r_, int s, bool count_ = true) : r(r_), stage(s), count(count_) {
    if (r->timing) (void)hipEventRecord(r->ev0, r->ctx->stream);
  }
  ~StageTimer() {
    if (!r->timing) return;
    (void)hipEventRecord(r->ev1, r->ctx->stream);
    (void)hipEventSynchronize(r->ev1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, r->ev0, r->ev1);
    r->stage_ms[stage] += ms;
    if (count) r->stage_n[stage] += 1;
  }
};

// measurements the hard projection reads: half storage (Yh), the x-transformed
// copy in hybrid mode (Yx), or the caller's k-space data (Y)
const void* meas_buffer(const llr_recon* r) {
  if (r->s16) return r->Yh.p;
  return r->hybrid ? r->Yx.p : r->Y.p;
}

// Momentum (extrapolation of the iterate entering the LLR step) in
// coefficient space: with the fused wavelet the LLR step sees the iterate only
// through its coefficients C = X S, and C is linear, so the extrapolated
// iterate x_k + b (x_k - x_{k-1}) has coefficients (1 + b) C(x_k) - b C(x_{k-1}).
// Cprev keeps C(x_{k-1}); the first step after a set-up only stores it.
llr_status apply_momentum(llr_recon* r, int mode, bool backproject) {
  if (!(r->momentum > 0) || mode == 2 || backproject) return LLR_OK;
  const size_t n = r->N * r->p.srank;
  if (!r->Cprev.p) LLR_HIP(r->Cprev.alloc(n * 8));
  llr_status st = momentum_device(r->ctx, r->Cb.as<float2>(), r->Cprev.as<float2>(), n, (float)r->momentum,
                                  r->mom_first);
  r->mom_first = false;
  return st;
}

// backproject = false leaves the new coefficients in Cb (with a subspace) for
// the fused wavelet to expand.
llr_status recon_llr(llr_recon* r, const size_t* shift, int mode, float lambda, unsigned* smax,
                     bool backproject = true) {
  llr_ctx* ctx = r->ctx;
  const llr_recon_params& p = r->p;
  const bool partial = p.partial_blocks != 0;
  long long nb = 1;
  for (int a = 0; a < 3; ++a) nb *= (long long)(partial ? (p.vol[a] + p.blk[a] - 1) / p.blk[a] : p.vol[a] / p.blk[a]);
  llr_status st;
  if (p.precision == LLR_PREC_MIXED) {
    // BF16 MFMA projection / Gram, FP32 refinement sweep, FP64 eigen-solve for ill-conditioned blocks
    unsigned* fp64n = r->counters.as<unsigned>();
    if (mode != 2) r->llr_blocks += (unsigned long long)nb;
    float2* F = p.srank ? r->Cb.as<float2>() : r->X.as<float2>();
    if (p.srank && !r->cb_adj) {
      StageTimer t(r, 1);
      st = r->s16 ? project_s16_device(ctx, r->X.p, F, r->S.as<float2>(), r->N, (int)r->nq, (int)p.srank, 1.f)
                  : project_bf16_device(ctx, r->X.as<float2>(), F, r->S.as<float2>(), r->N, (int)r->nq, (int)p.srank);
      if (st) return st;
    }
    // momentum (learned warm start) is applied by the block kernel as it loads C
    float2* prev = nullptr;
    const bool mom_first = r->mom_first;
    if (p.srank && r->momentum > 0 && mode != 2 && !backproject) {
      if (!r->Cprev.p) LLR_HIP(r->Cprev.alloc(r->N * p.srank * 8));
      prev = r->Cprev.as<float2>();
      r->mom_first = false;
    }
    {
      StageTimer t(r, 0);
      st = llr_block_run_mixed(ctx, F, p.vol, p.blk, shift, r->nc, 0, 0, nb, mode, (int)p.rank, lambda, smax, nullptr,
                               mode != 2 ? fp64n : nullptr, partial, prev, (float)r->momentum, mom_first);
    }
    if (st || mode == 2 || !p.srank || !backproject) return st;
    StageTimer t(r, 1);
    if (r->s16)
      return backproject_device(ctx, F, r->X.p, r->S.as<float2>(), p.vol, p.blk, shift, (int)r->nq, (int)p.srank,
                                partial, true, 1.f);
    return backproject_bf16_device(ctx, F, r->X.as<float2>(), r->S.as<float2>(), p.vol, p.blk, shift, (int)r->nq,
                                   (int)p.srank, partial);
  }
  if (p.srank) {
    if (!r->cb_adj) {
      StageTimer t(r, 1);
      st = project_device(ctx, r->X.as<float2>(), r->Cb.as<float2>(), r->S.as<float2>(), r->N, (int)r->nq,
                          (int)p.srank);
    }
    if (!st) st = apply_momentum(r, mode, backproject);
    if (st) return st;
    {
      // low-rank coefficients replace C in place; X = L S^H is formed by a
      // streaming pass in voxel order (blocks' voxels only)
      StageTimer t(r, 0);
      st = llr_block_run(ctx, r->Cb.as<float2>(), r->Cb.as<float2>(), nullptr, 0, p.vol, p.blk, shift, r->nc, 0, 0, nb,
                         mode, (int)p.rank, lambda, smax, nullptr, partial);
    }
    if (st || mode == 2 || !backproject) return st;
    StageTimer t(r, 1);
    return backproject_device(ctx, r->Cb.as<float2>(), r->X.as<float2>(), r->S.as<float2>(), p.vol, p.blk, shift,
                              (int)r->nq, (int)p.srank, partial);
  }
  StageTimer t(r, 0);
  if (r->nc > 8)
    return llr_block_run_wide(ctx, r->X.as<float2>(), p.vol, p.blk, shift, r->nc, 0, 0, nb, mode, (int)p.rank, lambda,
                              smax, partial, r->wide.as<float2>());
  return llr_block_run(ctx, r->X.as<float2>(), r->X.as<float2>(), nullptr, 0, p.vol, p.blk, shift, r->nc, 0, 0, nb, mode,
                       (int)p.rank, lambda, smax, nullptr, partial);
}

// Forward model -> hard projection with its FP64 residual into *err -> adjoint
// (lib.rs:311-333), in place on the iterate. The volumes are independent in
// these passes: they run in groups of r->vgroup volumes (all at once when 0),
// each group through all its passes before the next, so a group's
// intermediate data can stay in the Infinity Cache between the passes.
// to_coef (with r->yproj): the adjoint's inverse y pass writes the subspace
// coefficients of its output into Cb and X keeps that pass's input.
llr_status data_consistency(llr_recon* r, double* err, bool to_coef = false) {
  llr_ctx* ctx = r->ctx;
  const llr_recon_params& p = r->p;
  llr_status st = LLR_OK;
  const bool h = r->s16;
  const size_t eb = h ? 4 : 8;  // stored bytes per complex value (iterate and measurements alike)
  const RampTables& T = r->tabs;
  const float sx = (float)(1.0 / std::sqrt((double)p.vol[0])), sy = (float)(1.0 / std::sqrt((double)p.vol[1]));
  const size_t G = (r->vgroup && r->vgroup < r->nq) ? r->vgroup : r->nq;
  size_t np_total = 0;
  for (size_t g0 = 0; g0 < r->nq; g0 += G) {
    const size_t ng = std::min(G, r->nq - g0);
    const bool first = g0 == 0;
    void* X = (char*)r->X.p + g0 * r->N * eb;
    const void* Ym = (const char*)meas_buffer(r) + g0 * r->N * eb;
    const float2 *pk0 = T.pk[0].as<float2>() + g0 * p.vol[0], *pi0 = T.pi[0].as<float2>() + g0 * p.vol[0];
    const float2 *pk1 = T.pk[1].as<float2>() + g0 * p.vol[1], *pi1 = T.pi[1].as<float2>() + g0 * p.vol[1];
    const float2 *pk2 = T.pk[2].as<float2>() + g0 * p.vol[2], *pi2 = T.pi[2].as<float2>() + g0 * p.vol[2];
    if (!r->hybrid) {
      StageTimer t(r, 3, first);
      st = fft_axis_pass(ctx, X, X, p.vol, ng, 0, false, pk0, false, pi0, true, sx, h);
    }
    if (st) return st;
    {
      StageTimer t(r, 4, first);
      st = fft_axis_pass(ctx, X, X, p.vol, ng, 1, false, pk1, false, pi1, true, sy, h);
    }
    if (st) return st;
    {
      StageTimer t(r, 5, first);
      size_t np = 0;
      st = fft_zproj(ctx, X, Ym, p.vol, ng, pk2, pi2, r->partials.as<double>() + np_total, r->parts_cap - np_total,
                     &np, h);
      np_total += np;
    }
    if (st) return st;
    if (to_coef && r->yproj && ng == r->nq) {
      StageTimer t(r, 15);
      st = fft_y_adjoint_project(ctx, X, p.vol, r->nq, pi1, pk1, sy, r->Cb.as<float2>(), r->S.as<float2>(),
                                 (int)p.srank, r->s_real, h, T.pi_sign[1]);
      if (st) return st;
      r->cb_adj = true;
      r->x_pending = true;
      continue;
    }
    {
      StageTimer t(r, 9, first);
      st = fft_axis_pass(ctx, X, X, p.vol, ng, 1, true, pi1, false, pk1, true, sy, h);
    }
    if (st) return st;
    if (!r->hybrid) {
      StageTimer t(r, 10, first);
      st = fft_axis_pass(ctx, X, X, p.vol, ng, 0, true, pi0, false, pk0, true, sx, h);
    }
    if (st) return st;
  }
  StageTimer t(r, 6);
  return reduce_strided(ctx, r->partials.as<double>(), np_total, 1, err);
}


llr_status recon_iteration(llr_recon* r, size_t it, bool to_coef = false) {
  llr_ctx* ctx = r->ctx;
  const llr_recon_params& p = r->p;
  llr_status st = recon_llr(r, &r->shifts[3 * it], p.llr_mode, (float)r->lambda, nullptr, !r->fuse_wavelet);
  r->cb_adj = false;
  if (st) return st;
  if (p.wavelet_alpha > 0) {
    StageTimer t(r, 2);
    const float alpha = (float)(p.wavelet_alpha * r->hs);
    if (r->fuse_wavelet) {
      const WaveletFromSubspace sub{r->Cb.as<float2>(), r->S.as<float2>(), (int)p.srank, r->coefK.as<float2>(), r->s_real};
      st = wavelet_denoise_subspace(ctx, r->X.p, r->nq, p.vol, alpha, r->wL, r->coef.p, sub, r->s16);
      r->x_pending = false;  // every voxel of X rewritten from the coefficients
    } else {
      st = wavelet_denoise_device(ctx, r->X.p, r->nq, p.vol, alpha, r->wL, r->coef.p, r->t1.p, r->t2.p, r->s16);
    }
    if (st) return st;
  }
  if (r->x_pending) return fail(LLR_ERR_ARG, "internal: image not formed before the forward model");
  return data_consistency(r, r->hist.as<double>() + it, to_coef);
}

}  // namespace

extern "C" {

llr_status llr_recon_create(llr_ctx* ctx, const llr_recon_params* params, llr_recon** out) {
  if (!ctx || !params || !out) return fail(LLR_ERR_ARG, "null argument");
  const llr_recon_params& p = *params;
  llr_status st = check_fft_dims(p.vol);
  if (st) return st;
  if ((st = check_blocks(p.vol, p.blk))) return st;
  if (p.nq == 0) return fail(LLR_ERR_ARG, "no volumes");
  size_t m = p.blk[0] * p.blk[1] * p.blk[2];
  if (m <= p.nq) return fail(LLR_ERR_ARG, "only tall matrices are supported");  // lib.rs:203
  int nc = (int)(p.srank ? p.srank : p.nq);
  if (p.llr_mode == LLR_MODE_RANK && (p.rank == 0 || p.rank > (size_t)nc))
    return fail(LLR_ERR_ARG, "rank must be in 1..=columns");  // lib.rs:205-207
  if (p.srank && p.rank > p.srank && p.llr_mode == LLR_MODE_RANK)
    return fail(LLR_ERR_ARG, "llr rank must be less than or equal to subspace rank");  // lib.rs:217
  // more than 8 columns: the wide LLR kernels, for the FP32 form without a subspace (lib.rs:290-292)
  if (nc > 8 && (p.srank || p.precision != LLR_PREC_FP32))
    return fail(LLR_ERR_UNSUPPORTED, "more than 8 matrix columns need the FP32 form without a subspace");
  if (nc > 128) return fail(LLR_ERR_UNSUPPORTED, "the LLR kernels handle up to 128 matrix columns");
  if (p.srank && !p.subspace) return fail(LLR_ERR_ARG, "subspace rank given without a subspace matrix");
  if (p.n_iter && !p.shifts) return fail(LLR_ERR_ARG, "shift sequence required");
  if (p.precision != LLR_PREC_FP32 && p.precision != LLR_PREC_MIXED) return fail(LLR_ERR_ARG, "unknown precision form");
  if (p.precision == LLR_PREC_MIXED) {
    if (llr_block_mixed_smem((long long)(p.blk[0] * p.blk[1] * p.blk[2]), nc) > 64 * 1024)
      return fail(LLR_ERR_UNSUPPORTED, "block too large for the mixed-precision LLR kernel");
    if (p.srank && p.nq > 128) return fail(LLR_ERR_UNSUPPORTED, "mixed form handles at most 128 volumes");
  }
  if (p.storage != LLR_STORE_FP32 && p.storage != LLR_STORE_FP16) return fail(LLR_ERR_ARG, "unknown storage");
  if (p.storage == LLR_STORE_FP16 && (p.precision != LLR_PREC_MIXED || !p.srank))
    return fail(LLR_ERR_UNSUPPORTED, "half storage needs the mixed-precision form and a subspace");
  if (p.storage == LLR_STORE_FP16 && (p.vol[0] * p.vol[1] * p.vol[2] * p.nq) % 2)
    return fail(LLR_ERR_UNSUPPORTED, "half storage needs an even number of samples (nx*ny*nz*nq)");
  LLR_HIP(hipSetDevice(ctx->device));
  llr_recon* r = new llr_recon();
  r->ctx = ctx;
  r->p = p;
  r->nq = p.nq;
  r->N = p.vol[0] * p.vol[1] * p.vol[2];
  r->D = r->N * p.nq;
  r->nc = nc;
  r->s16 = p.storage == LLR_STORE_FP16;
  const size_t eb = r->s16 ? 4 : 8;  // bytes per stored complex value
  r->shifts.assign(p.shifts, p.shifts + 3 * p.n_iter);
  if (const char* g = std::getenv("LLR_VOLUME_GROUP")) r->vgroup = (size_t)std::strtoull(g, nullptr, 10);
  r->p.shifts = nullptr;
  r->p.subspace = nullptr;
  auto bail = [&](llr_status s) {
    delete r;
    return s;
  };
  if ((st = upload(r->X, nullptr, r->D * eb))) return bail(st);
  if ((st = upload(r->Y, nullptr, r->D * 8))) return bail(st);
  if (nc > 8) {  // Gram / W per block of the wide LLR kernels
    long long nb = 1;
    for (int a = 0; a < 3; ++a)
      nb *= (long long)(p.partial_blocks ? (p.vol[a] + p.blk[a] - 1) / p.blk[a] : p.vol[a] / p.blk[a]);
    if ((st = upload(r->wide, nullptr, llr_wide_work_elems(nc, nb) * 8))) return bail(st);
  }
  if (r->s16 && (st = upload(r->Yh, nullptr, r->D * 4))) return bail(st);
  if (p.srank) {
    if ((st = upload(r->Cb, nullptr, r->N * p.srank * 8))) return bail(st);
    if ((st = upload(r->S, p.subspace, p.nq * p.srank * 8))) return bail(st);
    // a real subspace (zero imaginary parts, as the log-linear DTI basis) takes
    // the wavelet's two-FMA expansion
    r->s_real = true;
    for (size_t i = 0; i < p.nq * p.srank; ++i)
      if (p.subspace[i].im != 0.f) r->s_real = false;
  }
  if (p.wavelet_alpha > 0) {
    r->wL = p.wavelet_levels < 0 ? wavelet_max_level(p.vol) : p.wavelet_levels;
    size_t ce = 0, te = 0;
    wavelet_workspace_elems(p.vol, r->wL, &ce, &te);
    if ((st = upload(r->coef, nullptr, ce * p.nq * eb))) return bail(st);
    bool cover = p.partial_blocks != 0;
    if (!cover) {
      cover = true;
      for (int a = 0; a < 3; ++a) cover = cover && p.vol[a] % p.blk[a] == 0;
    }
    r->fuse_wavelet = p.srank && cover && wavelet_subspace_supported(p.vol, r->wL, (int)p.srank);
    if (r->fuse_wavelet && (st = upload(r->coefK, nullptr, wavelet_subspace_elems(p.vol, r->wL, (int)p.srank) * 8)))
      return bail(st);
    if ((st = upload(r->t1, nullptr, te * p.nq * 8))) return bail(st);
    if ((st = upload(r->t2, nullptr, te * p.nq * 8))) return bail(st);
  }
  // the image ramps (lin_img_corr) are fixed for the reconstruction; pk is redone by every set-up
  if ((st = make_tables(r->tabs, p.vol, p.nq, nullptr, p.lin_img_corr))) return bail(st);
  r->p.lin_img_corr = nullptr;
  // partial sums: the fused z pass's, at any tiles-per-workgroup setting (the
  // batch autotuner may change it), or the set-up's energy/arg-max pass
  size_t np = 0;
  fft_zproj(ctx, nullptr, nullptr, p.vol, p.nq, nullptr, nullptr, nullptr, 0, &np, r->s16);
  r->parts_cap = std::max<size_t>(np, 2 * energy_argmax_parts(r->N, p.nq));
  if ((st = upload(r->partials, nullptr, r->parts_cap * 8))) return bail(st);
  if ((st = upload(r->stats, nullptr, 16))) return bail(st);
  if ((st = upload(r->hist, nullptr, std::max<size_t>(p.n_iter, 1) * 8))) return bail(st);
  if ((st = upload(r->smax, nullptr, 8))) return bail(st);
  if ((st = upload(r->best, nullptr, p.nq * 8))) return bail(st);
  if ((st = upload(r->counters, nullptr, 8))) return bail(st);
  if ((st = upload(r->werr, nullptr, 8))) return bail(st);
  LLR_HIP(hipMemset(r->counters.p, 0, 8));
  if (hipEventCreate(&r->ev0) != hipSuccess || hipEventCreate(&r->ev1) != hipSuccess)
    return bail(fail(LLR_ERR_HIP, "hipEventCreate failed"));
  *out = r;
  return LLR_OK;
}

void llr_recon_destroy(llr_recon* r) {
  if (!r) return;
  (void)hipSetDevice(r->ctx->device);
  (void)hipStreamSynchronize(r->ctx->stream);
  if (r->ctx->kt_ms == r->stage_ms) {
    r->ctx->kt_ms = nullptr;
    r->ctx->kt_n = nullptr;
  }
  if (r->ev0) (void)hipEventDestroy(r->ev0);
  if (r->ev1) (void)hipEventDestroy(r->ev1);
  delete r;
}

llr_status llr_recon_set_measurements(llr_recon* r, const llr_cf32* y, size_t len, int on_device) {
  if (!r || !y) return fail(LLR_ERR_ARG, "null argument");
  if (len != r->D) return fail(LLR_ERR_ARG, "measurements must hold nx*ny*nz*nq samples");
  LLR_HIP(hipSetDevice(r->ctx->device));
  LLR_HIP(hipMemcpyAsync(r->Y.p, y, len * 8, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                         r->ctx->stream));
  LLR_HIP(hipStreamSynchronize(r->ctx->stream));
  r->have_meas = true;
  return LLR_OK;
}

llr_status llr_recon_set_initial(llr_recon* r, const llr_cf32* x, size_t len, int on_device) {
  if (!r || !x) return fail(LLR_ERR_ARG, "null argument");
  if (len != r->D) return fail(LLR_ERR_ARG, "initial image must hold nx*ny*nz*nq samples");
  LLR_HIP(hipSetDevice(r->ctx->device));
  if (r->s16 && !r->Xin.p) {
    llr_status st = upload(r->Xin, nullptr, r->D * 8);
    if (st) return st;
  }
  LLR_HIP(hipMemcpyAsync(r->s16 ? r->Xin.p : r->X.p, x, len * 8,
                         on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, r->ctx->stream));
  // Half storage holds values at the scale of the last set-up. Before a set-up
  // the start waits in complex64 and the set-up converts it at its scale;
  // after llr_recon_prepare (the next run skips the set-up) it is converted now
  // at the scale in force.
  r->xin_pending = r->s16 && !r->prepared;
  if (r->s16 && r->prepared) {
    llr_status st = to_h16_device(r->ctx, r->Xin.as<float2>(), r->X.p, r->D, (float)r->hs);
    if (st) return st;
  }
  LLR_HIP(hipStreamSynchronize(r->ctx->stream));
  r->have_init = true;
  return LLR_OK;
}

}  // extern "C"

namespace llr {

// Set-up of a run (lib.rs:224-258): total energy, linear k-space corrections
// and ramp tables, zero-filled adjoint (unless warm-started), SVT threshold.
static llr_status recon_setup(llr_recon* r) {
  llr_ctx* ctx = r->ctx;
  const llr_recon_params& p = r->p;
  llr_status st;
  r->cb_current = false;
  r->cb_adj = false;
  r->x_pending = false;
  // total energy (lib.rs:224, 409-417) and the arg-max behind the linear
  // k-space corrections (lib.rs:229-231) in one pass over the measurements
  {
    StageTimer t(r, 7);
    LLR_HIP(hipMemsetAsync(r->best.p, 0, r->nq * 8, ctx->stream));
    if ((st = energy_argmax_device(ctx, r->Y.as<float2>(), r->N, r->nq, r->partials.as<double>(),
                                   r->best.as<unsigned long long>())))
      return st;
    if ((st = reduce_strided(ctx, r->partials.as<double>(), energy_argmax_parts(r->N, r->nq), 2,
                             r->stats.as<double>())))
      return st;
  }
  // the k-space ramp tables follow from the arg-max on the device; the image
  // ramps (zero corrections) were made when the reconstruction was created
  for (int a = 0; a < 3; ++a)
    if ((st = ramp_from_argmax_device(ctx, r->tabs.pk[a].as<float2>(), p.vol, a, r->nq,
                                      r->best.as<unsigned long long>())))
      return st;
  std::vector<unsigned long long> best(r->nq);
  double stats[2];
  LLR_HIP(hipStreamSynchronize(ctx->stream));
  LLR_HIP(hipMemcpy(best.data(), r->best.p, r->nq * 8, hipMemcpyDeviceToHost));
  LLR_HIP(hipMemcpy(stats, r->stats.p, 16, hipMemcpyDeviceToHost));
  r->energy = stats[0];
  r->hs = (r->s16 && r->energy > 0) ? std::sqrt((double)r->N / r->energy) : 1.0;
  r->ksp_corr.assign(3 * r->nq, 0.f);
  for (size_t q = 0; q < r->nq; ++q) {
    size_t idx = (size_t)(0xFFFFFFFFu - (unsigned)(best[q] & 0xFFFFFFFFull));
    size_t s[3] = {idx % p.vol[0], (idx / p.vol[0]) % p.vol[1], idx / (p.vol[0] * p.vol[1])};
    for (int a = 0; a < 3; ++a) r->ksp_corr[3 * q + a] = (float)((long long)s[a] - ((long long)p.vol[a] + 1) / 2);
  }
  // The measurements the iteration reads. Unless the caller asks for the full
  // 3D iteration, they go through the inverse x pass once (into half storage
  // directly when the iterate is held in half), which also checks that every
  // x-line is fully measured or empty before and after (as stored): the hard
  // projection's measured-sample test then picks whole lines, and the x
  // transform commutes with it.
  r->hybrid = false;
  {
    StageTimer t(r, 7);
    const float sx = (float)(1.0 / std::sqrt((double)p.vol[0]));
    const RampTables& T = r->tabs;
    bool convert = r->s16;
    if (!p.no_hybrid) {
      if (!r->flags.p && (st = upload(r->flags, nullptr, 8))) return st;
      if (!r->s16 && !r->Yx.p && (st = upload(r->Yx, nullptr, r->D * 8))) return st;
      LLR_HIP(hipMemsetAsync(r->flags.p, 0, 8, ctx->stream));
      if ((st = fft_x_inverse_checked(ctx, r->Y.as<float2>(), r->s16 ? r->Yh.p : r->Yx.p, r->s16, p.vol, r->nq,
                                      T.pi[0].as<float2>(), T.pk[0].as<float2>(),
                                      (float)(sx * (r->s16 ? r->hs : 1.0)), r->flags.as<unsigned>())))
        return st;
      unsigned bad = 0;
      LLR_HIP(hipStreamSynchronize(ctx->stream));
      LLR_HIP(hipMemcpy(&bad, r->flags.p, 4, hipMemcpyDeviceToHost));
      r->hybrid = bad == 0;
      convert = r->s16 && !r->hybrid;
    }
    // automatic (0): complex64 storage only; at 256x256x128x67 the fused pass
    // takes 1.64 ms against 1.65 + 1.01 ms for the y pass and the projection in
    // the iteration, but with half storage 1.60 ms against 0.85 + 0.77 ms
    r->yproj = (p.fused_projection > 0 || (p.fused_projection == 0 && !r->s16)) && r->fuse_wavelet && r->hybrid &&
               p.srank && fft_yproj_supported(p.vol[1], (int)p.srank) && (r->vgroup == 0 || r->vgroup >= r->nq);
    if (convert && (st = to_h16_device(ctx, r->Y.as<float2>(), r->Yh.p, r->D, (float)r->hs, true))) return st;
    // the initial image, if one was set
    if (r->s16 && r->have_init && r->xin_pending &&
        (st = to_h16_device(ctx, r->Xin.as<float2>(), r->X.p, r->D, (float)r->hs)))
      return st;
    r->xin_pending = false;
  }
  // zero-filled adjoint (lib.rs:233-240) unless warm-started
  if (!r->have_init) {
    StageTimer t(r, 7);
    if ((st = model_device(ctx, meas_buffer(r), r->X.p, p.vol, r->nq, r->tabs, LLR_INVERSE, r->s16,
                           r->hybrid ? 1 : 0)))
      return st;
  }
  // threshold: lambda_factor * sigma_max of the starting blocks at shift 0
  if (p.llr_mode == LLR_MODE_SVT) {
    if (p.svt_lambda > 0) {
      r->lambda = p.svt_lambda * r->hs;  // thresholds act on stored (scaled) values
    } else {
      LLR_HIP(hipMemsetAsync(r->smax.p, 0, 8, ctx->stream));
      size_t zero[3] = {0, 0, 0};
      if ((st = recon_llr(r, zero, 2, 0.f, r->smax.as<unsigned>()))) return st;
      // singular values only: the blocks are not written back, and the
      // projection is project_current's unless the mixed form rounds X to BF16
      r->cb_current = p.srank && (p.precision != LLR_PREC_MIXED || r->s16);
      unsigned bits = 0;
      LLR_HIP(hipStreamSynchronize(ctx->stream));
      LLR_HIP(hipMemcpy(&bits, r->smax.p, 4, hipMemcpyDeviceToHost));
      float smax;
      std::memcpy(&smax, &bits, 4);
      r->lambda = p.lambda_factor * (double)smax;
    }
  }
  return LLR_OK;
}

// C = X S of the current iterate (FP32 VALU from complex64, BF16 MFMA from half storage)
// (out_scale multiplies C: 1 / hs gives the caller's units from half storage)
static llr_status project_current(llr_recon* r, float2* C, float out_scale = 1.f) {
  if (r->s16)
    return project_s16_device(r->ctx, r->X.p, C, r->S.as<float2>(), r->N, (int)r->nq, (int)r->p.srank, out_scale);
  return project_device(r->ctx, r->X.as<float2>(), C, r->S.as<float2>(), r->N, (int)r->nq, (int)r->p.srank);
}

// X = scale * C S^H for every voxel (one block covering the whole volume)
static llr_status backproject_all(llr_recon* r, const float2* C, float scale = 1.f) {
  const size_t zero[3] = {0, 0, 0};
  return backproject_device(r->ctx, C, r->X.p, r->S.as<float2>(), r->p.vol, r->p.vol, zero, (int)r->nq,
                            (int)r->p.srank, false, r->s16, scale);
}

// X holds the adjoint's inverse y pass input (the fused adjoint projection ran):
// that pass, as the unfused adjoint runs it, forms the image.
static llr_status form_pending_image(llr_recon* r) {
  if (!r->x_pending) return LLR_OK;
  const llr_recon_params& p = r->p;
  const float sy = (float)(1.0 / std::sqrt((double)p.vol[1]));
  StageTimer t(r, 9);
  llr_status st = fft_axis_pass(r->ctx, r->X.p, r->X.p, p.vol, r->nq, 1, true, r->tabs.pi[1].as<float2>(), false,
                                r->tabs.pk[1].as<float2>(), true, sy, r->s16);
  r->x_pending = false;
  r->cb_adj = false;
  return st;
}

// Iterations it = 0 .. n_max-1 of the prepared run; stop (if > 0) ends the run
// once the relative residual falls to that value. Relative residuals out.
static llr_status run_iterations(llr_recon* r, size_t n_max, double stop, double* err_hist, size_t* iters_done,
                                 size_t min_iter = 0) {
  llr_ctx* ctx = r->ctx;
  llr_status st;
  size_t it = 0;
  for (; it < n_max; ++it) {
    // the last iteration's adjoint forms the image (no coefficients needed)
    if ((st = recon_iteration(r, it, it + 1 < n_max))) {
      r->x_pending = r->cb_adj = false;
      return st;
    }
    if (stop > 0 && it + 1 >= min_iter) {
      double e = 0;
      LLR_HIP(hipStreamSynchronize(ctx->stream));
      LLR_HIP(hipMemcpy(&e, r->hist.as<double>() + it, 8, hipMemcpyDeviceToHost));
      if (e / (r->energy * r->hs * r->hs) <= stop) {
        ++it;
        break;
      }
    }
  }
  if ((st = form_pending_image(r))) return st;
  LLR_HIP(hipStreamSynchronize(ctx->stream));
  if (err_hist && it) {
    std::vector<double> h(it);
    LLR_HIP(hipMemcpy(h.data(), r->hist.p, it * 8, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < it; ++i) err_hist[i] = h[i] / (r->energy * r->hs * r->hs);
  }
  if (iters_done) *iters_done = it;
  return LLR_OK;
}

// Set-up from the zero-filled adjoint (threshold included), then the iterate
// becomes the U-Net's prediction from the adjoint's subspace coefficients,
// made data-consistent (forward model, hard projection, adjoint) like every
// iterate the loop hands to the LLR step. residual_out: the prediction's
// relative data-consistency error before the projection.
static llr_status warm_start(llr_recon* r, llr_unet* net, const float* to_net, const float* from_net,
                             double* scale_out, double* residual_out) {
  llr_status st;
  r->have_init = false;
  if ((st = recon_setup(r))) return st;
  const int K = (int)r->p.srank;
  float2* C = r->Cb.as<float2>();
  if (!r->cb_current) {
    StageTimer t(r, 1);
    if ((st = project_current(r, C))) return st;
  }
  double scale = 0;
  {
    StageTimer t(r, 8);
    if ((st = unet_coef_scale(r->ctx, C, K, r->N, to_net, &scale))) return st;
    if (!(scale > 0) || !std::isfinite(scale)) return fail(LLR_ERR_ARG, "zero-filled adjoint has no signal");
    llr_ctx* nctx = unet_context(net);
    if (nctx != r->ctx) LLR_HIP(hipStreamSynchronize(r->ctx->stream));
    if ((st = unet_predict_device(net, C, K, r->p.vol, (float)scale, to_net, from_net, C))) return st;
    if (nctx != r->ctx) LLR_HIP(hipStreamSynchronize(nctx->stream));
  }
  {
    StageTimer t(r, 1);
    if ((st = backproject_all(r, C))) return st;
  }
  if ((st = data_consistency(r, r->werr.as<double>()))) return st;
  r->have_init = true;
  r->prepared = true;
  if (scale_out) *scale_out = scale;
  if (residual_out) {
    double e = 0;
    LLR_HIP(hipStreamSynchronize(r->ctx->stream));
    LLR_HIP(hipMemcpy(&e, r->werr.p, 8, hipMemcpyDeviceToHost));
    *residual_out = e / (r->energy * r->hs * r->hs);
  }
  return LLR_OK;
}

}  // namespace llr

extern "C" {

llr_status llr_recon_run(llr_recon* r, double* err_hist, size_t* iters_done) {
  if (!r) return fail(LLR_ERR_ARG, "null recon");
  if (!r->have_meas) return fail(LLR_ERR_ARG, "measurements not set");
  LLR_HIP(hipSetDevice(r->ctx->device));
  llr_status st;
  if (!r->prepared && (st = recon_setup(r))) return st;
  r->prepared = false;
  r->have_init = false;
  return run_iterations(r, r->p.n_iter, r->p.residual_stop, err_hist, iters_done);
}

llr_status llr_recon_prepare(llr_recon* r) {
  if (!r) return fail(LLR_ERR_ARG, "null recon");
  if (!r->have_meas) return fail(LLR_ERR_ARG, "measurements not set");
  LLR_HIP(hipSetDevice(r->ctx->device));
  llr_status st = recon_setup(r);
  if (st) return st;
  LLR_HIP(hipStreamSynchronize(r->ctx->stream));
  r->prepared = true;
  return LLR_OK;
}

llr_status llr_recon_get_coefficients(llr_recon* r, llr_cf32* out, size_t len, int on_device) {
  if (!r || !out) return fail(LLR_ERR_ARG, "null argument");
  if (!r->p.srank) return fail(LLR_ERR_ARG, "the reconstruction has no subspace");
  if (len != r->N * r->p.srank) return fail(LLR_ERR_ARG, "output must hold nx*ny*nz*srank samples");
  LLR_HIP(hipSetDevice(r->ctx->device));
  float2* dst = on_device ? (float2*)out : r->Cb.as<float2>();
  llr_status st = project_current(r, dst, (float)(1.0 / r->hs));
  if (st) return st;
  if (!on_device) return download(out, r->Cb, len * 8, r->ctx);
  LLR_HIP(hipStreamSynchronize(r->ctx->stream));
  return LLR_OK;
}

llr_status llr_recon_set_coefficients(llr_recon* r, const llr_cf32* c, size_t len, int on_device) {
  if (!r || !c) return fail(LLR_ERR_ARG, "null argument");
  if (!r->p.srank) return fail(LLR_ERR_ARG, "the reconstruction has no subspace");
  if (len != r->N * r->p.srank) return fail(LLR_ERR_ARG, "coefficients must hold nx*ny*nz*srank samples");
  LLR_HIP(hipSetDevice(r->ctx->device));
  const float2* src = (const float2*)c;
  if (!on_device) {
    LLR_HIP(hipMemcpyAsync(r->Cb.p, c, len * 8, hipMemcpyHostToDevice, r->ctx->stream));
    src = r->Cb.as<float2>();
  }
  llr_status st;
  if (r->s16 && !r->prepared) {
    // half storage before a set-up: the scale is not known yet, so the start
    // C S^H waits in complex64 and the set-up converts it at its scale (as
    // llr_recon_set_initial)
    if (!r->Xin.p && (st = upload(r->Xin, nullptr, r->D * 8))) return st;
    const size_t zero[3] = {0, 0, 0};
    st = backproject_device(r->ctx, src, r->Xin.p, r->S.as<float2>(), r->p.vol, r->p.vol, zero, (int)r->nq,
                            (int)r->p.srank, false, false, 1.f);
    r->xin_pending = true;
  } else {
    st = backproject_all(r, src, (float)r->hs);
    r->xin_pending = false;
  }
  if (st) return st;
  LLR_HIP(hipStreamSynchronize(r->ctx->stream));
  r->have_init = true;
  return LLR_OK;
}

llr_status llr_recon_warm_start(llr_recon* r, llr_unet* net, const float* to_net, const float* from_net,
                                double* scale, double* residual) {
  if (!r || !net) return fail(LLR_ERR_ARG, "null argument");
  if (!r->p.srank) return fail(LLR_ERR_ARG, "the warm start works in subspace-coefficient space: no subspace");
  if (!r->have_meas) return fail(LLR_ERR_ARG, "measurements not set");
  LLR_HIP(hipSetDevice(r->ctx->device));
  llr_status st = warm_start(r, net, to_net, from_net, scale, residual);
  if (st) return st;
  LLR_HIP(hipStreamSynchronize(r->ctx->stream));
  return LLR_OK;
}

llr_status llr_recon_run_learned(llr_recon* r, llr_unet* net, const llr_warm_params* wp, double* err_hist,
                                 size_t* iters_done, int* fell_back) {
  if (!r || !net || !wp) return fail(LLR_ERR_ARG, "null argument");
  if (!r->p.srank) return fail(LLR_ERR_ARG, "the warm start works in subspace-coefficient space: no subspace");
  if (!r->have_meas) return fail(LLR_ERR_ARG, "measurements not set");
  if (!(wp->residual_stop > 0)) return fail(LLR_ERR_ARG, "residual_stop must be positive");
  const size_t budget = wp->max_iter ? std::min(wp->max_iter, r->p.n_iter) : r->p.n_iter;
  LLR_HIP(hipSetDevice(r->ctx->device));
  if (wp->momentum < 0 || wp->momentum >= 1) return fail(LLR_ERR_ARG, "momentum must be in [0, 1)");
  if (wp->momentum > 0 && !r->fuse_wavelet)
    return fail(LLR_ERR_UNSUPPORTED, "momentum needs the wavelet fused with the subspace back-projection");
  llr_status st = warm_start(r, net, wp->to_net, wp->from_net, nullptr, nullptr);
  if (st) return st;
  size_t done = 0;
  std::vector<double> h(std::max<size_t>(r->p.n_iter, 1));
  r->momentum = wp->momentum;
  r->mom_first = true;
  st = run_iterations(r, budget, wp->residual_stop, h.data(), &done, wp->min_iter);
  r->momentum = 0.0;
  if (st) return st;
  const bool reached = done > 0 && h[done - 1] <= wp->residual_stop;
  if (fell_back) *fell_back = reached ? 0 : 1;
  if (!reached) {
    // cold start: the plain run from the zero-filled adjoint, n_iter iterations
    r->have_init = false;
    if ((st = recon_setup(r))) return st;
    if ((st = run_iterations(r, r->p.n_iter, 0.0, h.data(), &done))) return st;
  }
  r->prepared = false;
  r->have_init = false;
  if (err_hist)
    for (size_t i = 0; i < done; ++i) err_hist[i] = h[i];
  if (iters_done) *iters_done = done;
  return LLR_OK;
}

llr_status llr_recon_get_image(llr_recon* r, llr_cf32* out, size_t len, int on_device) {
  if (!r || !out) return fail(LLR_ERR_ARG, "null argument");
  if (len != r->D) return fail(LLR_ERR_ARG, "output must hold nx*ny*nz*nq samples");
  LLR_HIP(hipSetDevice(r->ctx->device));
  if (r->s16) {
    // widen to complex64: straight into a device output, else through a
    // temporary device buffer
    if (on_device) {
      llr_status st = from_h16_device(r->ctx, r->X.p, (float2*)out, len, (float)(1.0 / r->hs));
      if (st) return st;
    } else {
      DevBuf tmp;
      LLR_HIP(tmp.alloc(len * 8));
      llr_status st = from_h16_device(r->ctx, r->X.p, tmp.as<float2>(), len, (float)(1.0 / r->hs));
      if (st) return st;
      LLR_HIP(hipMemcpyAsync(out, tmp.p, len * 8, hipMemcpyDeviceToHost, r->ctx->stream));
      LLR_HIP(hipStreamSynchronize(r->ctx->stream));
    }
    LLR_HIP(hipStreamSynchronize(r->ctx->stream));
    return LLR_OK;
  }
  LLR_HIP(hipMemcpyAsync(out, r->X.p, len * 8, on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                         r->ctx->stream));
  LLR_HIP(hipStreamSynchronize(r->ctx->stream));
  return LLR_OK;
}

llr_status llr_recon_info(llr_recon* r, double* lambda, float* ksp_corr) {
  if (!r) return fail(LLR_ERR_ARG, "null recon");
  if (lambda) *lambda = r->lambda / r->hs;
  if (ksp_corr && !r->ksp_corr.empty()) std::memcpy(ksp_corr, r->ksp_corr.data(), r->ksp_corr.size() * sizeof(float));
  return LLR_OK;
}

llr_status llr_recon_hybrid(llr_recon* r, int* hybrid) {
  if (!r || !hybrid) return fail(LLR_ERR_ARG, "null argument");
  *hybrid = r->hybrid ? 1 : 0;
  return LLR_OK;
}

llr_status llr_recon_mixed_stats(llr_recon* r, unsigned long long* fp64_blocks, unsigned long long* blocks) {
  if (!r) return fail(LLR_ERR_ARG, "null recon");
  LLR_HIP(hipSetDevice(r->ctx->device));
  LLR_HIP(hipStreamSynchronize(r->ctx->stream));
  unsigned n = 0;
  LLR_HIP(hipMemcpy(&n, r->counters.p, sizeof(unsigned), hipMemcpyDeviceToHost));
  if (fp64_blocks) *fp64_blocks = n;
  if (blocks) *blocks = r->llr_blocks;
  return LLR_OK;
}

llr_status llr_recon_enable_timing(llr_recon* r, int enable) {
  if (!r) return fail(LLR_ERR_ARG, "null recon");
  r->timing = enable != 0;
  llr_ctx* c = r->ctx;
  if (r->timing && !c->kt_ev[0]) {
    LLR_HIP(hipSetDevice(c->device));
    LLR_HIP(hipEventCreate(&c->kt_ev[0]));
    LLR_HIP(hipEventCreate(&c->kt_ev[1]));
  }
  c->kt_ms = r->timing ? r->stage_ms : nullptr;
  c->kt_n = r->timing ? r->stage_n : nullptr;
  for (int i = 0; i < LLR_NUM_STAGES; ++i) {
    r->stage_ms[i] = 0;
    r->stage_n[i] = 0;
  }
  return LLR_OK;
}

llr_status llr_recon_stage_times(llr_recon* r, double* total_ms, size_t* launches) {
  if (!r) return fail(LLR_ERR_ARG, "null recon");
  for (int i = 0; i < LLR_NUM_STAGES; ++i) {
    if (total_ms) total_ms[i] = r->stage_ms[i];
    if (launches) launches[i] = r->stage_n[i];
  }
  return LLR_OK;
}

}  // extern "C"

// ======================================================= batched reconstructions
// Many reconstructions of uneven size run together: each member keeps its own
// buffers and stream; one iteration of all members is captured once as a HIP
// graph (fork per member: projection -> join: one persistent LLR launch over
// the blocks of every member -> fork: back-projection, wavelet, forward model,
// hard projection, adjoint, residual -> join: iteration counter += 1) and the
// graph is replayed n_iter times. Every per-iteration quantity (the block
// shift, the residual slot) is read on the device from the iteration counter.
struct llr_recon_batch {
  llr_ctx* ctx = nullptr;               // caller's context (device)
  hipStream_t s0 = nullptr;             // capture / replay stream
  std::vector<llr_ctx*> mctx;           // one private context (stream, tables) per member
  std::vector<llr_recon*> rec;
  std::vector<DevBuf*> shifts, tmp;     // device shift tables (int n_iter x 3), reduction scratch
  DevBuf iter, tune_buf;
  MultiPlan plan;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev_fork = nullptr, ev_mid = nullptr;
  std::vector<hipEvent_t> ev_a, ev_b;
  int autotune = 0;
  bool tuned = false;
  size_t n_iter = 0;
  std::vector<int> stage;               // LLR LDS staging per member (-1 auto)
};

namespace llr {

static const size_t kBlockCandidates[3] = {6, 8, 10};

// Time `fn` (launches on stream s) as the minimum of `reps` runs, in ms.
template <class F>
static llr_status time_min(hipStream_t s, int reps, F fn, float* ms_out) {
  hipEvent_t a, b;
  LLR_HIP(hipEventCreate(&a));
  LLR_HIP(hipEventCreate(&b));
  float best = 1e30f;
  llr_status st = LLR_OK;
  for (int i = 0; i < reps && !st; ++i) {
    (void)hipEventRecord(a, s);
    st = fn();
    (void)hipEventRecord(b, s);
    if (!st && hipEventSynchronize(b) == hipSuccess) {
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, a, b);
      best = std::min(best, ms);
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *ms_out = best;
  return st;
}

// Autotuner, per member, after its set-up: block size from {6,8,10}^3 (those
// that fit the volume with more rows than columns), by the time of one LLR
// step per voxel; LDS staging of the block kernel; FFT tiles per workgroup per
// axis from {1, 2, 3, 4, 6, 8}. The LLR step is timed on the coefficient planes
// (subspace) or in singular-value-only mode (no subspace), so the iterate is
// not modified; FFT passes are timed out of place into a scratch buffer.
static llr_status batch_autotune(llr_recon_batch* B) {
  for (size_t j = 0; j < B->rec.size(); ++j) {
    llr_recon* r = B->rec[j];
    llr_ctx* ctx = r->ctx;
    llr_recon_params& p = r->p;
    const size_t zero[3] = {0, 0, 0};
    float best_t = 1e30f;
    size_t best_b = p.blk[0];
    double best_lam = r->lambda;
    for (size_t bc : kBlockCandidates) {
      if (bc > p.vol[0] || bc > p.vol[1] || bc > p.vol[2] || bc * bc * bc <= (size_t)r->nc) continue;
      size_t blk[3] = {bc, bc, bc};
      for (int a = 0; a < 3; ++a) p.blk[a] = bc;
      double lam = p.svt_lambda;
      if (p.llr_mode == LLR_MODE_SVT && p.svt_lambda <= 0) {
        LLR_HIP(hipMemsetAsync(r->smax.p, 0, 8, ctx->stream));
        llr_status st = recon_llr(r, zero, 2, 0.f, r->smax.as<unsigned>());
        if (st) return st;
        unsigned bits = 0;
        LLR_HIP(hipStreamSynchronize(ctx->stream));
        LLR_HIP(hipMemcpy(&bits, r->smax.p, 4, hipMemcpyDeviceToHost));
        float smax;
        std::memcpy(&smax, &bits, 4);
        lam = p.lambda_factor * (double)smax;
      }
      long long nb = 1;
      for (int a = 0; a < 3; ++a) nb *= (long long)((p.vol[a] + bc - 1) / bc);
      float ms = 0.f;
      llr_status st = time_min(ctx->stream, 2, [&]() -> llr_status {
        if (p.srank) {
          llr_status s2 = project_device(ctx, r->X.as<float2>(), r->Cb.as<float2>(), r->S.as<float2>(), r->N,
                                         (int)r->nq, (int)p.srank);
          if (s2) return s2;
          return llr_block_run(ctx, r->Cb.as<float2>(), r->Cb.as<float2>(), nullptr, 0, p.vol, blk, zero, r->nc, 0,
                               0, nb, p.llr_mode, (int)p.rank, (float)lam, nullptr, nullptr, true);
        }
        return llr_block_run(ctx, r->X.as<float2>(), r->X.as<float2>(), nullptr, 0, p.vol, blk, zero, r->nc, 0, 0,
                             nb, 2, 0, 0.f, nullptr, nullptr, true);
      }, &ms);
      if (st) return st;
      if (ms < best_t * 0.97f) {  // prefer the earlier (smaller) block unless clearly slower
        best_t = ms;
        best_b = bc;
        best_lam = lam;
      }
    }
    for (int a = 0; a < 3; ++a) p.blk[a] = best_b;
    r->lambda = best_lam;
    // LDS staging of the block kernel for the chosen size
    {
      size_t blk[3] = {best_b, best_b, best_b};
      long long nb = 1;
      for (int a = 0; a < 3; ++a) nb *= (long long)((p.vol[a] + best_b - 1) / best_b);
      float t_on = 0.f, t_off = 0.f;
      for (int stg = 1; stg >= 0; --stg) {
        ctx->llr_stage = stg;
        llr_status st = time_min(ctx->stream, 2, [&]() -> llr_status {
          return llr_block_run(ctx, p.srank ? r->Cb.as<float2>() : r->X.as<float2>(),
                               p.srank ? r->Cb.as<float2>() : r->X.as<float2>(), nullptr, 0, p.vol, blk, zero, r->nc,
                               0, 0, nb, 2, 0, 0.f, nullptr, nullptr, true);
        }, stg ? &t_on : &t_off);
        if (st) return st;
      }
      ctx->llr_stage = (t_off < 0.9f * t_on) ? 0 : 1;
      B->stage[j] = ctx->llr_stage;
    }
    // FFT tiles per workgroup, per axis (out of place into the scratch buffer)
    const RampTables& T = r->tabs;
    for (int a = 0; a < 3; ++a) {
      float bt = 1e30f;
      int bw = 4;
      for (int w : {1, 2, 3, 4, 6, 8}) {
        ctx->fft_tpw[a] = w;
        float ms = 0.f;
        float sc = (float)(1.0 / std::sqrt((double)p.vol[a]));
        llr_status st = time_min(ctx->stream, 2, [&]() -> llr_status {
          return fft_axis_pass(ctx, r->X.as<float2>(), B->tune_buf.as<float2>(), p.vol, r->nq, a, false,
                               T.pk[a].as<float2>(), false, T.pi[a].as<float2>(), true, sc);
        }, &ms);
        if (st) return st;
        if (ms < bt * 0.97f) {
          bt = ms;
          bw = w;
        }
      }
      ctx->fft_tpw[a] = bw;
    }
  }
  B->tuned = true;
  return LLR_OK;
}

// Work of member j before the LLR launch (projection onto the subspace; with
// the fused adjoint projection the previous replay's adjoint, or the
// projection made before the first replay, has left the coefficients).
static llr_status batch_pre(llr_recon* r) {
  if (!r->p.srank || r->yproj) return LLR_OK;
  return project_device(r->ctx, r->X.as<float2>(), r->Cb.as<float2>(), r->S.as<float2>(), r->N, (int)r->nq,
                        (int)r->p.srank);
}

// Work of member j after the LLR launch: back-projection, wavelet, forward
// model, hard projection + residual, adjoint, residual into hist[*iter].
static llr_status batch_post(llr_recon* r, const unsigned* iter, double* tmp) {
  llr_ctx* ctx = r->ctx;
  const llr_recon_params& p = r->p;
  const size_t zero[3] = {0, 0, 0};
  llr_status st;
  if (p.srank && !r->fuse_wavelet &&
      (st = backproject_device(ctx, r->Cb.as<float2>(), r->X.as<float2>(), r->S.as<float2>(), p.vol, p.blk, zero,
                               (int)r->nq, (int)p.srank, true)))
    return st;
  if (p.wavelet_alpha > 0) {
    if (r->fuse_wavelet) {
      // the members' grids have partial edge blocks: every voxel is covered
      const WaveletFromSubspace sub{r->Cb.as<float2>(), r->S.as<float2>(), (int)p.srank, r->coefK.as<float2>(), r->s_real};
      st = wavelet_denoise_subspace(ctx, r->X.p, r->nq, p.vol, (float)p.wavelet_alpha, r->wL, r->coef.p, sub, false);
    } else {
      st = wavelet_denoise_device(ctx, r->X.p, r->nq, p.vol, (float)p.wavelet_alpha, r->wL, r->coef.p, r->t1.p,
                                  r->t2.p);
    }
    if (st) return st;
  }
  float2* X = r->X.as<float2>();
  const RampTables& T = r->tabs;
  float sx = (float)(1.0 / std::sqrt((double)p.vol[0])), sy = (float)(1.0 / std::sqrt((double)p.vol[1]));
  if (!r->hybrid && (st = fft_axis_pass(ctx, X, X, p.vol, r->nq, 0, false, T.pk[0].as<float2>(), false,
                                        T.pi[0].as<float2>(), true, sx)))
    return st;
  if ((st = fft_axis_pass(ctx, X, X, p.vol, r->nq, 1, false, T.pk[1].as<float2>(), false, T.pi[1].as<float2>(), true,
                          sy)))
    return st;
  size_t np = 0;
  if ((st = fft_zproj(ctx, X, meas_buffer(r), p.vol, r->nq, T.pk[2].as<float2>(), T.pi[2].as<float2>(),
                      r->partials.as<double>(), r->parts_cap, &np)))
    return st;
  if (r->yproj) {
    if ((st = fft_y_adjoint_project(ctx, X, p.vol, r->nq, T.pi[1].as<float2>(), T.pk[1].as<float2>(), sy,
                                    r->Cb.as<float2>(), r->S.as<float2>(), (int)p.srank, r->s_real, false,
                                    T.pi_sign[1])))
      return st;
  } else if ((st = fft_axis_pass(ctx, X, X, p.vol, r->nq, 1, true, T.pi[1].as<float2>(), false,
                                 T.pk[1].as<float2>(), true, sy))) {
    return st;
  }
  if (!r->hybrid && (st = fft_axis_pass(ctx, X, X, p.vol, r->nq, 0, true, T.pi[0].as<float2>(), false,
                                        T.pk[0].as<float2>(), true, sx)))
    return st;
  return reduce_at(ctx, r->partials.as<double>(), np, r->hist.as<double>(), iter, tmp);
}

static llr_status batch_capture(llr_recon_batch* B) {
  if (B->exec) {
    (void)hipGraphExecDestroy(B->exec);
    B->exec = nullptr;
  }
  if (B->graph) {
    (void)hipGraphDestroy(B->graph);
    B->graph = nullptr;
  }
  LLR_HIP(hipStreamBeginCapture(B->s0, hipStreamCaptureModeRelaxed));
  llr_status st = LLR_OK;
  const size_t n = B->rec.size();
  unsigned* iter = B->iter.as<unsigned>();
  if (hipEventRecord(B->ev_fork, B->s0) != hipSuccess) st = fail(LLR_ERR_HIP, "event record in capture");
  for (size_t j = 0; j < n && !st; ++j) {
    hipStream_t sj = B->mctx[j]->stream;
    if (hipStreamWaitEvent(sj, B->ev_fork, 0) != hipSuccess) st = fail(LLR_ERR_HIP, "stream wait in capture");
    if (!st) st = batch_pre(B->rec[j]);
    if (!st && (hipEventRecord(B->ev_a[j], sj) != hipSuccess || hipStreamWaitEvent(B->s0, B->ev_a[j], 0) != hipSuccess))
      st = fail(LLR_ERR_HIP, "join in capture");
  }
  if (!st) st = llr_multi_launch(B->s0, B->plan, iter);
  if (!st && hipEventRecord(B->ev_mid, B->s0) != hipSuccess) st = fail(LLR_ERR_HIP, "event record in capture");
  for (size_t j = 0; j < n && !st; ++j) {
    hipStream_t sj = B->mctx[j]->stream;
    if (hipStreamWaitEvent(sj, B->ev_mid, 0) != hipSuccess) st = fail(LLR_ERR_HIP, "stream wait in capture");
    if (!st) st = batch_post(B->rec[j], iter, B->tmp[j]->as<double>());
    if (!st && (hipEventRecord(B->ev_b[j], sj) != hipSuccess || hipStreamWaitEvent(B->s0, B->ev_b[j], 0) != hipSuccess))
      st = fail(LLR_ERR_HIP, "join in capture");
  }
  if (!st) st = advance_counter(B->s0, iter);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(B->s0, &g);
  if (st) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  B->graph = g;
  LLR_HIP(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
  return LLR_OK;
}

}  // namespace llr

extern "C" {

void llr_recon_batch_destroy(llr_recon_batch* B) {
  if (!B) return;
  (void)hipSetDevice(B->ctx->device);
  if (B->s0) (void)hipStreamSynchronize(B->s0);
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  for (llr_recon* r : B->rec) llr_recon_destroy(r);
  for (llr_ctx* c : B->mctx) llr_ctx_destroy(c);
  for (DevBuf* d : B->shifts) delete d;
  for (DevBuf* d : B->tmp) delete d;
  for (hipEvent_t e : B->ev_a) (void)hipEventDestroy(e);
  for (hipEvent_t e : B->ev_b) (void)hipEventDestroy(e);
  if (B->ev_fork) (void)hipEventDestroy(B->ev_fork);
  if (B->ev_mid) (void)hipEventDestroy(B->ev_mid);
  if (B->s0) (void)hipStreamDestroy(B->s0);
  delete B;
}

llr_status llr_recon_batch_create(llr_ctx* ctx, const llr_recon_params* params, size_t n, int autotune,
                                  llr_recon_batch** out) {
  if (!ctx || !params || !out || n == 0) return fail(LLR_ERR_ARG, "null argument or empty batch");
  LLR_HIP(hipSetDevice(ctx->device));
  const size_t n_iter = params[0].n_iter;
  for (size_t j = 0; j < n; ++j) {
    if (params[j].n_iter != n_iter) return fail(LLR_ERR_ARG, "batch members must share n_iter");
    if (params[j].residual_stop > 0) return fail(LLR_ERR_ARG, "residual_stop is not available in a batch");
    if (params[j].precision != LLR_PREC_FP32) return fail(LLR_ERR_UNSUPPORTED, "batches run the FP32 form");
  }
  llr_recon_batch* B = new llr_recon_batch();
  B->ctx = ctx;
  B->autotune = autotune;
  B->n_iter = n_iter;
  auto bail = [&](llr_status s) {
    llr_recon_batch_destroy(B);
    return s;
  };
  if (hipStreamCreateWithFlags(&B->s0, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&B->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&B->ev_mid, hipEventDisableTiming) != hipSuccess)
    return bail(fail(LLR_ERR_HIP, "stream/event creation failed"));
  llr_status st;
  size_t maxD = 0;
  for (size_t j = 0; j < n; ++j) {
    llr_ctx* c = nullptr;
    if ((st = llr_ctx_create(ctx->device, &c))) return bail(st);
    B->mctx.push_back(c);
    llr_recon_params p = params[j];
    p.partial_blocks = 1;  // uneven sizes: edge blocks are partial
    llr_recon* r = nullptr;
    if ((st = llr_recon_create(c, &p, &r))) return bail(st);
    B->rec.push_back(r);
    std::vector<int> sh(3 * n_iter);
    for (size_t i = 0; i < 3 * n_iter; ++i) sh[i] = (int)(p.shifts[i] % p.vol[i % 3]);
    DevBuf* d = new DevBuf();
    B->shifts.push_back(d);
    if ((st = upload(*d, sh.data(), std::max<size_t>(sh.size(), 3) * sizeof(int)))) return bail(st);
    DevBuf* t = new DevBuf();
    B->tmp.push_back(t);
    if ((st = upload(*t, nullptr, 256 * sizeof(double)))) return bail(st);
    hipEvent_t a, b;
    if (hipEventCreateWithFlags(&a, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b, hipEventDisableTiming) != hipSuccess)
      return bail(fail(LLR_ERR_HIP, "event creation failed"));
    B->ev_a.push_back(a);
    B->ev_b.push_back(b);
    maxD = std::max(maxD, r->D);
  }
  B->stage.assign(n, -1);
  if ((st = upload(B->iter, nullptr, 16))) return bail(st);
  if (autotune && (st = upload(B->tune_buf, nullptr, maxD * 8))) return bail(st);
  *out = B;
  return LLR_OK;
}

llr_status llr_recon_batch_set_measurements(llr_recon_batch* B, size_t i, const llr_cf32* y, size_t len,
                                            int on_device) {
  if (!B || i >= B->rec.size()) return fail(LLR_ERR_ARG, "no such batch member");
  return llr_recon_set_measurements(B->rec[i], y, len, on_device);
}

llr_status llr_recon_batch_run(llr_recon_batch* B, double* err_hist) {
  if (!B) return fail(LLR_ERR_ARG, "null batch");
  LLR_HIP(hipSetDevice(B->ctx->device));
  llr_status st;
  for (llr_recon* r : B->rec) {
    if (!r->have_meas) return fail(LLR_ERR_ARG, "measurements not set for every member");
    if ((st = recon_setup(r))) return st;
  }
  if (B->autotune && !B->tuned && (st = batch_autotune(B))) return st;
  std::vector<MemberSpec> ms(B->rec.size());
  for (size_t j = 0; j < B->rec.size(); ++j) {
    llr_recon* r = B->rec[j];
    const llr_recon_params& p = r->p;
    MemberSpec& m = ms[j];
    m.F = p.srank ? r->Cb.as<float2>() : r->X.as<float2>();
    for (int a = 0; a < 3; ++a) {
      m.vol[a] = p.vol[a];
      m.blk[a] = p.blk[a];
    }
    m.nc = r->nc;
    m.mode = p.llr_mode;
    m.rank = (int)p.rank;
    m.lambda = (float)r->lambda;
    m.d_shifts = B->shifts[j]->as<int>();
    m.partial = true;
    m.stage = B->stage[j];
  }
  if ((st = llr_multi_plan(B->plan, ms))) return st;
  for (llr_recon* r : B->rec) LLR_HIP(hipStreamSynchronize(r->ctx->stream));
  if ((st = batch_capture(B))) return st;
  // members with the fused adjoint projection: the first LLR step's
  // coefficients come from the set-up's image, projected once here
  if (B->n_iter)
    for (llr_recon* r : B->rec)
      if (r->yproj) {
        if ((st = project_current(r, r->Cb.as<float2>()))) return st;
        LLR_HIP(hipStreamSynchronize(r->ctx->stream));
      }
  LLR_HIP(hipMemsetAsync(B->iter.p, 0, 16, B->s0));  // outside the graph
  for (size_t it = 0; it < B->n_iter; ++it) LLR_HIP(hipGraphLaunch(B->exec, B->s0));
  LLR_HIP(hipStreamSynchronize(B->s0));
  // the last replay's adjoint left its input in X: form the images
  if (B->n_iter)
    for (llr_recon* r : B->rec)
      if (r->yproj) {
        r->x_pending = true;
        if ((st = form_pending_image(r))) return st;
        LLR_HIP(hipStreamSynchronize(r->ctx->stream));
      }
  if (err_hist) {
    for (size_t j = 0; j < B->rec.size(); ++j) {
      llr_recon* r = B->rec[j];
      std::vector<double> h(B->n_iter);
      if (B->n_iter) LLR_HIP(hipMemcpy(h.data(), r->hist.p, B->n_iter * 8, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < B->n_iter; ++i) err_hist[j * B->n_iter + i] = h[i] / r->energy;
    }
  }
  return LLR_OK;
}

llr_status llr_recon_batch_get_image(llr_recon_batch* B, size_t i, llr_cf32* out, size_t len, int on_device) {
  if (!B || i >= B->rec.size()) return fail(LLR_ERR_ARG, "no such batch member");
  return llr_recon_get_image(B->rec[i], out, len, on_device);
}

llr_status llr_recon_batch_tuning(llr_recon_batch* B, size_t i, size_t blk[3], int fft_tpw[3], int* llr_stage,
                                  double* lambda) {
  if (!B || i >= B->rec.size()) return fail(LLR_ERR_ARG, "no such batch member");
  llr_recon* r = B->rec[i];
  for (int a = 0; a < 3; ++a) {
    if (blk) blk[a] = r->p.blk[a];
    if (fft_tpw) fft_tpw[a] = r->ctx->fft_tpw[a];
  }
  if (llr_stage) *llr_stage = r->ctx->llr_stage;
  if (lambda) *lambda = r->lambda;
  return LLR_OK;
}

}  // extern "C"
