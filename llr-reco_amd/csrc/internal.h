// Internal interfaces between the kernel files and the C ABI layer.
#pragma once

#include <array>
#include <vector>

#include "common.h"

namespace llr {

// fft.hip
bool fft_size_supported(size_t n);
// src/dst/data/meas hold complex64, or half-precision pairs (ch16) when h16 is set
// (ramped signal-model passes only).
// twiddles (exp(-2 pi i m / n), m < n) may be given; else the context's cached table
llr_status fft_axis_pass(llr_ctx* ctx, const void* src, void* dst, const size_t* dims, size_t nvols, int axis,
                         bool inverse, const float2* pre, bool pre_conj, const float2* post, bool post_conj,
                         float scale, bool h16 = false, const float2* twiddles = nullptr);
// Inverse x pass (ramps pre, conj(post), times scale) of complex64 src into
// dst (complex64, or half-precision pairs when h16_out), setting *flag to 1 if
// some x-line of src or of the stored dst is partly zero (neither all nonzero
// nor all zero). *flag is not cleared.
llr_status fft_x_inverse_checked(llr_ctx* ctx, const float2* src, void* dst, bool h16_out, const size_t* dims,
                                 size_t nvols, const float2* pre, const float2* post, float scale, unsigned* flag);
// Fused forward-z / hard projection / adjoint-z: one partial error sum per
// workgroup into partial (cap doubles; the launch fails if it would need
// more). partial == null queries the largest count any tiles-per-workgroup
// setting of ctx can produce (the buffer size to allocate).
llr_status fft_zproj(llr_ctx* ctx, void* data, const void* meas, const size_t* dims, size_t nvols,
                     const float2* pk, const float2* pi, double* partial, size_t cap, size_t* nparts,
                     bool h16 = false);
// The adjoint's inverse y pass fused with the subspace projection: C[k][n] =
// sum_q X[q][n] S[q][k] (K planes of N voxels) of the pass's output X, which is
// not stored (src keeps the pass's input). y lengths 16..256 that are 2^a 3^b
// with radices of at least 4, K <= 8. pre_sign: the input ramps are the parity
// pattern (-1)^(i + ny/2) (taken as exact signs; pre is not read).
bool fft_yproj_supported(size_t ny, int K);
llr_status fft_y_adjoint_project(llr_ctx* ctx, const void* src, const size_t* dims, size_t nvols, const float2* pre,
                                 const float2* post, float scale, float2* C, const float2* S, int K, bool s_real,
                                 bool h16, bool pre_sign = false);

// llr_wide.hip: LLR prox of blocks with more than 8 columns (up to 128).
// work: llr_wide_work_elems(nc, nblocks) complex values, or null to allocate.
size_t llr_wide_work_elems(int nc, long long nblocks);
llr_status llr_block_run_wide(llr_ctx* ctx, float2* F, const size_t* vol, const size_t* blk, const size_t* shift,
                              int nc, int layout, long long block0, long long nblocks, int mode, int rank,
                              float lambda, unsigned* smax_bits, bool partial, float2* work);

// llr.hip
// One reconstruction's share of the persistent multi-reconstruction LLR launch.
struct MemberSpec {
  float2* F;            // nc planes, processed in place
  size_t vol[3], blk[3];
  int nc, mode, rank;
  float lambda;
  const int* d_shifts;  // device n_iter x 3 shift table
  bool partial;
  int stage;            // LDS staging override (-1 auto)
};
struct MultiPlan {
  DevBuf desc;
  int nm = 0;
  long long total = 0;
  size_t smem = 0;
  long long per_wg = 0;
  int per_xcd = 0;
  unsigned grid = 0;
};
llr_status llr_multi_plan(MultiPlan& plan, const std::vector<MemberSpec>& members);
// `iter` points at the device iteration counter that selects each member's shift
llr_status llr_multi_launch(hipStream_t s, const MultiPlan& plan, const unsigned* iter);
llr_status lift_device(llr_ctx* ctx, const void* vols, void* mats, size_t nvols, size_t eb, const size_t* vol,
                       const size_t* blk, const size_t* shift, size_t start, size_t n, bool unlift);
llr_status project_device(llr_ctx* ctx, const float2* X, float2* Cp, const float2* S, size_t N, int nq, int K);
// X[q][v] = sum_k C[k][v] conj(S[q][k]) for voxels inside the block grid of `shift`
// (X as complex64, or half-precision pairs when h16 is set)
llr_status backproject_device(llr_ctx* ctx, const float2* C, void* X, const float2* S, const size_t* vol,
                              const size_t* blk, const size_t* shift, int nq, int K, bool partial = false,
                              bool h16 = false, float scale = 1.f);
llr_status llr_block_run(llr_ctx* ctx, const float2* F, float2* out, const float2* S, int nq, const size_t* vol,
                         const size_t* blk, const size_t* shift, int nc, int layout, long long block0,
                         long long nblocks, int mode, int rank, float lambda, unsigned* smax_bits, float* svals,
                         bool partial = false);
llr_status lowrank_generic_device(llr_ctx* ctx, float2* mats, long long m, int n, long long batch, int mode, int rank,
                                  float lambda, unsigned* smax_bits);

// wavelet.hip
int wavelet_max_level(const size_t* dims);
void wavelet_geometry(const size_t* dims, int L, std::vector<std::array<size_t, 3>>& in_dims,
                      std::vector<std::array<size_t, 3>>& cdims);
size_t wavelet_workspace_elems(const size_t* dims, int L, size_t* coef_elems, size_t* tmp_elems);
llr_status project_bf16_device(llr_ctx* ctx, const float2* X, float2* Cp, const float2* S, size_t N, int nq, int K);
// C = X S on BF16 MFMA (hi + lo split of each half value) for an iterate held in
// half storage (ch16), times out_scale
llr_status project_s16_device(llr_ctx* ctx, const void* X, float2* Cp, const float2* S, size_t N, int nq, int K,
                              float out_scale);
llr_status backproject_bf16_device(llr_ctx* ctx, const float2* C, float2* X, const float2* S, const size_t* vol,
                                   const size_t* blk, const size_t* shift, int nq, int K, bool partial);
// prev (optional, layout 0 with a block grid covering every voxel): the
// momentum extrapolation F <- (1 + mom) F - mom prev is applied as each block is
// loaded and prev <- F stored (only stored when mom_first).
llr_status llr_block_run_mixed(llr_ctx* ctx, float2* F, const size_t* vol, const size_t* blk, const size_t* shift,
                               int nc, int layout, long long block0, long long nblocks, int mode, int rank,
                               float lambda, unsigned* smax_bits, float* svals, unsigned* fp64_count, bool partial,
                               float2* prev = nullptr, float mom = 0.f, bool mom_first = false);
size_t llr_block_mixed_smem(long long m, int nc);
llr_status dti_scalars_device(llr_ctx* ctx, const float2* vols, size_t n_vox, size_t nq, const float* bvals,
                              const float* bvecs, float b0_tol, float* out);
// The wavelet of an iterate given by subspace coefficients: X_q = sum_k C_k
// conj(S[q][k]) at every voxel (K planes C, FP32). Every analysis runs on the
// K planes; BK is their FP32 workspace of wavelet_subspace_elems values.
struct WaveletFromSubspace {
  const float2* C;
  const float2* S;
  int K;
  float2* BK;
  bool s_real;  // S has zero imaginary parts: two FMAs per expansion term
};
size_t wavelet_subspace_elems(const size_t* dims, int L, int K);
bool wavelet_subspace_supported(const size_t* dims, int L, int K);
// The wavelet soft threshold of X_q = C S^H written into vols (every voxel).
llr_status wavelet_denoise_subspace(llr_ctx* ctx, void* vols, size_t nvols, const size_t* dims, float alpha, int L,
                                    void* coef, const WaveletFromSubspace& sub, bool h16);
// vols and coef hold complex64, or half-precision pairs (ch16) when h16 is set
llr_status wavelet_denoise_device(llr_ctx* ctx, void* vols, size_t nvols, const size_t* dims, float alpha, int L,
                                  void* coef, void* t1, void* t2, bool h16 = false);

// unet.hip
llr_status unet_predict_device(llr_unet* U, const float2* coef, int K, const size_t* vol, float scale,
                               const float* to_net, const float* from_net, float2* out);
llr_status unet_coef_scale(llr_ctx* ctx, const float2* coef, int K, size_t N, const float* to_net, double* scale);
llr_ctx* unet_context(llr_unet* U);

// misc.hip
llr_status shrink_all_device(llr_ctx* ctx, float2* x, size_t n, float alpha);
// k-space ramp table of one axis (nvols x dims[axis]) from the per-volume
// arg-max found by argmax_volumes (integer corrections, exact reduction)
llr_status ramp_from_argmax_device(llr_ctx* ctx, float2* tab, const size_t* dims, int axis, size_t nvols,
                                   const unsigned long long* best);
// complex64 -> half-precision pairs (x s) and back (x s); n complex values, n even.
// keep_nonzero: a nonzero value never becomes zero (measurements: the hard
// projection's measured set is the set of nonzero samples)
llr_status to_h16_device(llr_ctx* ctx, const float2* src, void* dst, size_t n, float s, bool keep_nonzero = false);
llr_status from_h16_device(llr_ctx* ctx, const void* src, float2* dst, size_t n, float s);
// C <- (1 + beta) C - beta P, P <- the old C (first: P <- C only); n complex values, n even
llr_status momentum_device(llr_ctx* ctx, float2* C, float2* P, size_t n, float beta, bool first);
llr_status energy_partials(llr_ctx* ctx, const float2* y, size_t n, double* partials, size_t nparts);
llr_status reduce_strided(llr_ctx* ctx, const double* partials, size_t n, size_t stride, double* out);
llr_status reduce_at(llr_ctx* ctx, const double* partials, size_t n, double* out, const unsigned* at, double* tmp);
llr_status advance_counter(hipStream_t s, unsigned* c);
llr_status reduce_sum(llr_ctx* ctx, const double* partials, size_t nparts, double* out, double* tmp, size_t tmp_n);
llr_status hard_project_partials(llr_ctx* ctx, float2* x, const float2* y, size_t n, double* partials, size_t nparts);
llr_status argmax_volumes(llr_ctx* ctx, const float2* y, size_t N, size_t nvols, unsigned long long* best);
// energy partials (sum |y|^2, nonzero count; 2 per part) and per-volume arg-max
// keys in one pass; energy_argmax_parts(N, nvols) parts; best zeroed by the caller
size_t energy_argmax_parts(size_t N, size_t nvols);
llr_status energy_argmax_device(llr_ctx* ctx, const float2* y, size_t N, size_t nvols, double* partials,
                                unsigned long long* best);
llr_status under_sample_device(llr_ctx* ctx, float2* data, const size_t* dims, size_t nq, const float2* mask);
llr_status circshift_device(llr_ctx* ctx, const void* src, void* dst, size_t eb, const size_t* dims, const long long* sh);
llr_status phase_shift3_device(llr_ctx* ctx, float2* x, const size_t* dims, int sign);
llr_status scale_device(llr_ctx* ctx, float2* x, size_t n, float s);
llr_status cdgmm_device(llr_ctx* ctx, float2* a, const float* d, size_t m, size_t n, size_t batch);
llr_status cgemm_device(llr_ctx* ctx, size_t m, size_t k, size_t n, size_t batch, bool bcast_a, float2 alpha,
                        float2 beta, bool ta, bool tb, const float2* A, const float2* B, float2* C);
llr_status conj_transpose_device(llr_ctx* ctx, const float2* src, float2* dst, size_t m, size_t n, size_t batch,
                                 bool conj);

// phase tables: exp(+i * 2*pi * frac((i - n/2) * (c + n/2) / n)) for each volume
// (rows) and index i; host-side construction in long double.
void ramp_table(std::vector<float2>& out, size_t n, const float* corr, size_t stride, size_t nvols);

}  // namespace llr
