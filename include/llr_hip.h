/*
 * llr_hip.h - C ABI of the MI355X-native LLR diffusion-MRI reconstruction path.
 *
 * Drop-in boundary for the CUDA path of wyatt-A/llr-reco. Every entry point
 * below replaces one Rust function of that crate (cited file:line) and keeps
 * its argument meaning, order and preconditions. Plain pointers and sizes only;
 * no exception or panic crosses the boundary: each call returns an llr_status
 * and llr_last_error() holds the message of the last failure on the calling
 * thread (the reference panics with the same conditions).
 *
 * Data conventions (shared with the reference):
 *   - complex64 = interleaved {re, im} float pairs (num_complex::Complex32 ==
 *     cuComplex, bindings.rs:9189-9191);
 *   - every array is column-major: voxel index x + nx*(y + ny*z), the volume
 *     index slowest (array_utils.rs:4-17, lib.rs:374-388);
 *   - a set of volumes is one contiguous buffer of n_volumes * nx*ny*nz
 *     elements (the slices lib.rs:374-388 builds are views of one such buffer).
 *
 * Host-pointer entry points are synchronous and allocate their device buffers
 * per call, as the reference wrappers do (e.g. dti_subspace.rs:150-165). The
 * device-resident reconstruction (llr_recon_*) keeps the iterate, the
 * measurements and every workspace in HBM for the whole run.
 *
 * Threading: a context is used by one host thread at a time. All device work of
 * a context runs on its own stream (default priority).
 */
#ifndef LLR_HIP_H
#define LLR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct llr_ctx llr_ctx;
typedef struct llr_recon llr_recon;

typedef enum {
  LLR_OK = 0,
  LLR_ERR_ARG = 1,         /* precondition violated (the reference asserts/panics) */
  LLR_ERR_HIP = 2,         /* HIP runtime failure */
  LLR_ERR_NOMEM = 3,       /* device allocation failed */
  LLR_ERR_UNSUPPORTED = 4, /* shape outside what the kernels implement */
  LLR_ERR_NUMERIC = 5      /* non-finite value met where the reference would fail */
} llr_status;

typedef struct {
  float re, im;
} llr_cf32;

/* Message of the last failing call on this thread ("" if none). */
const char* llr_last_error(void);
/* ABI version of this header (major*100 + minor). */
int llr_abi_version(void);

/* ---------------------------------------------------------------- context
 * Replaces the per-call device selection and allocation of cuda_api.rs:8-191:
 * one context owns a HIP stream and cached twiddle/phase tables. */
llr_status llr_ctx_create(int device, llr_ctx** out);
void llr_ctx_destroy(llr_ctx* ctx);
llr_status llr_ctx_synchronize(llr_ctx* ctx);
/* Device number of the context (as given to llr_ctx_create). */
int llr_ctx_device(const llr_ctx* ctx);

/* ---------------------------------------------------------------- block.rs */
/* grid_dim (block.rs:112-118): grid[a] = vol_size[a] / block_size[a]. */
void llr_grid_dim(const size_t vol_size[3], const size_t block_size[3], size_t grid[3]);

/* lift_data_set (block.rs:147-164): gather n_matrices shifted blocks
 * (block indices matrix_idx_start .. +n_matrices, corners from block_coord,
 * block.rs:122-133, periodic wrap block.rs:8-28) of n_volumes volumes into
 * column-major (prod(block_size) x n_volumes) matrices, contiguous.
 * elem_bytes is the size of the element type T (the Rust API is generic).
 * matrix_len (elements) must be >= prod(block_size)*n_volumes*n_matrices
 * (block.rs:152); every block dimension must be <= the volume dimension
 * (block.rs:69-71). Bit-exact data movement. */
llr_status llr_lift_data_set(llr_ctx* ctx, const void* volumes, size_t n_volumes, size_t elem_bytes,
                             void* matrix_data, size_t matrix_len, const size_t vol_size[3],
                             const size_t block_size[3], const size_t shift[3], size_t matrix_idx_start,
                             size_t n_matrices);

/* un_lift_data_set (block.rs:166-184): the inverse scatter of llr_lift_data_set. */
llr_status llr_un_lift_data_set(llr_ctx* ctx, void* volumes, size_t n_volumes, size_t elem_bytes,
                                const void* matrix_data, size_t matrix_len, const size_t vol_size[3],
                                const size_t block_size[3], const size_t shift[3], size_t matrix_idx_start,
                                size_t n_matrices);

/* lift_data_set / un_lift_data_set on buffers already on the context's device
 * (volumes and matrix_data are device pointers; same arguments and checks;
 * returns after the copy has finished). No counterpart in the reference, whose
 * buffers are host slices: for callers that keep the data resident. */
llr_status llr_lift_data_set_device(llr_ctx* ctx, const void* volumes, size_t n_volumes, size_t elem_bytes,
                                    void* matrix_data, size_t matrix_len, const size_t vol_size[3],
                                    const size_t block_size[3], const size_t shift[3], size_t matrix_idx_start,
                                    size_t n_matrices);
llr_status llr_un_lift_data_set_device(llr_ctx* ctx, void* volumes, size_t n_volumes, size_t elem_bytes,
                                       const void* matrix_data, size_t matrix_len, const size_t vol_size[3],
                                       const size_t block_size[3], const size_t shift[3], size_t matrix_idx_start,
                                       size_t n_matrices);

/* --------------------------------------------------------- dti_subspace.rs */
/* dti_project (dti_subspace.rs:141-166): compressed (srank x nvox per matrix)
 * = S^H X^H for each of batch_size image matrices X (nvox x nq); S is nq x srank.
 * compressed_len must equal nvox*srank*batch_size (dti_subspace.rs:148). */
llr_status llr_dti_project(llr_ctx* ctx, const llr_cf32* subspace, size_t subspace_len,
                           const llr_cf32* image_matrices, size_t image_len, llr_cf32* compressed,
                           size_t compressed_len, size_t batch_size, size_t nq, size_t nvox, size_t srank);

/* dti_back_project (dti_subspace.rs:172-207): image matrices (nvox x nq) =
 * (S * compressed)^H with compressed given as srank x nvox per matrix.
 * image_len must equal nq*nvox*batch_size (dti_subspace.rs:175). */
llr_status llr_dti_back_project(llr_ctx* ctx, const llr_cf32* subspace, size_t subspace_len,
                                llr_cf32* image_matrices, size_t image_len, const llr_cf32* compressed,
                                size_t compressed_len, size_t batch_size, size_t nq, size_t nvox,
                                size_t srank);

/* conj_transpose_matrix (dti_subspace.rs:211-221): (m x n) -> (n x m)^H per matrix. */
llr_status llr_conj_transpose_matrix(llr_ctx* ctx, llr_cf32* matrix_data, size_t len, size_t m, size_t n,
                                     size_t batch_size);

/* subspace_constraint (dti_subspace.rs:89-134): X <- transpose(S S^H X^H),
 * keeping the reference's final plain (non-conjugating) transpose. */
llr_status llr_subspace_constraint(llr_ctx* ctx, const llr_cf32* subspace, size_t subspace_len,
                                   llr_cf32* image_matrices, size_t image_len, size_t batch_size, size_t nq,
                                   size_t nvox, size_t srank);

/* ------------------------------------------------ low_rank.rs / cublas.rs */
/* cu_low_rank_approx_batch (low_rank.rs:10-69): replace each m x n matrix
 * (m > n, n <= 128) by its rank-`rank` truncation U_r diag(s_r) V_r^H.
 * n <= 8: one fused kernel per matrix; 8 < n <= 128: Gram, workgroup Jacobi
 * and apply kernels (one workgroup per matrix each). */
llr_status llr_low_rank_approx_batch(llr_ctx* ctx, size_t m, size_t n, size_t rank, size_t batch_size,
                                     llr_cf32* matrix_data, size_t len);

/* Singular-value soft threshold of each m x n matrix: U diag(max(s-lambda,0)) V^H.
 * No reference counterpart (the reference only truncates); same layout as
 * llr_low_rank_approx_batch. sigma_max (may be NULL) receives the largest
 * singular value met. */
llr_status llr_svt_batch(llr_ctx* ctx, size_t m, size_t n, float lambda, size_t batch_size,
                         llr_cf32* matrix_data, size_t len, float* sigma_max);

/* cublas_cgemm_strided_batched_device (cublas.rs:12-105) on host buffers:
 * C = alpha*op(A)*op(B) + beta*C, op = conjugate transpose when flagged, A
 * broadcast over the batch when broadcast_a != 0. */
llr_status llr_cgemm_strided_batched(llr_ctx* ctx, size_t m, size_t k, size_t n, size_t batch_size,
                                     int broadcast_a, llr_cf32 alpha, llr_cf32 beta, int transpose_a,
                                     int transpose_b, const llr_cf32* a, size_t a_len, const llr_cf32* b,
                                     size_t b_len, llr_cf32* c, size_t c_len);

/* rhs_cdgmm_batched_device (cublas.rs:119-127, kernels.cu:11-67): A[:,j,b] *= d[j,b]. */
llr_status llr_rhs_cdgmm_batched(llr_ctx* ctx, size_t m, size_t n, size_t batch_size, llr_cf32* matrix,
                                 size_t len, const float* diag);

/* ------------------------------------------------ signal_model.rs / cufft.rs */
#define LLR_FORWARD 0
#define LLR_INVERSE 1
#define LLR_NORM_NONE 0
#define LLR_NORM_INVERSE 1
#define LLR_NORM_UNITARY 2

/* signal_model (signal_model.rs:34-79): Forward = ramp(ksp corr, +1) -> unitary
 * forward DFT -> ramp(img corr, -1); Inverse = ramp(img, +1) -> unitary inverse
 * DFT -> ramp(ksp, -1); ramp phase (i - n/2) * 2*pi*(c + n/2)/n per axis
 * (cufft.rs:290-326, signal_model.rs:110-117). lin_*_corr hold batch_size x 3
 * floats. The ramps are evaluated exactly (turn-reduced), not in f32. */
llr_status llr_signal_model(llr_ctx* ctx, llr_cf32* x, size_t len, const size_t image_size[3],
                            size_t batch_size, const float* lin_ksp_corr, const float* lin_img_corr,
                            int direction);

/* Device-pointer form of llr_signal_model for callers that keep the data in HBM:
 * x is a device pointer on the context's device; returns when the transform is
 * complete. Same semantics and preconditions as llr_signal_model. */
llr_status llr_signal_model_device(llr_ctx* ctx, llr_cf32* x, size_t len, const size_t image_size[3],
                                   size_t batch_size, const float* lin_ksp_corr, const float* lin_img_corr,
                                   int direction);

/* llr_signal_model_device without any device allocation. No counterpart in
 * the reference; for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc): the twiddle and
 * phase-ramp tables are built in a fixed 2 MiB device array of the library
 * (262144 complex values, one per GPU, kept until the process ends) instead of
 * buffers allocated per call. Same arguments, semantics and completion as
 * llr_signal_model_device; sizes whose tables need more than
 * batch_size * (2 (nx+ny+nz) + 3) + nx+ny+nz <= 262144 entries are rejected
 * with LLR_ERR_UNSUPPORTED. Calls that overlap on one GPU (several threads or
 * contexts) share the array and race. */
llr_status llr_signal_model_device_noalloc(llr_ctx* ctx, llr_cf32* x, size_t len, const size_t image_size[3],
                                           size_t batch_size, const float* lin_ksp_corr, const float* lin_img_corr,
                                           int direction);

/* estimate_linear_correction (signal_model.rs:100-106 via max_coord,
 * cufft.rs:329-337): arg-max of |k|^2 (lowest index on ties) minus (n+1)/2. */
llr_status llr_estimate_linear_correction(llr_ctx* ctx, const llr_cf32* ksp, size_t len,
                                          const size_t vol_size[3], float out[3]);

/* cu_fftn_batch (cufft.rs:124-180): batched in-place 3-D C2C transform of
 * col-major volumes, norm = LLR_NORM_* (cufft.rs:19-29, 152-166). */
llr_status llr_fftn_batch(llr_ctx* ctx, const size_t dims[3], size_t batches, int direction, int norm,
                          llr_cf32* data, size_t len);

/* Centered unitary 3-D FFT of the batch_size col-major volumes in x, in place
 * (fft3c / ifft3c / fft3c_chunked / ifft3c_chunked, cufft.rs:340-402):
 * phase_shift3(Inverse) -> unitary FFT (LLR_FORWARD or LLR_INVERSE) ->
 * phase_shift3(Forward), the FOV/2 modulation pi (i - n/2) per axis evaluated
 * exactly. vols_per_batch > 0 moves that many volumes through the device at
 * a time (the chunked forms); 0 moves all at once. */
llr_status llr_fft3c(llr_ctx* ctx, llr_cf32* x, size_t len, const size_t dims[3], size_t batch_size, int direction,
                     size_t vols_per_batch);
/* cu_fft3 (cufft.rs:213-267): one volume, unnormalised forward, 1/N on the inverse. */
llr_status llr_cu_fft3(llr_ctx* ctx, llr_cf32* x, size_t len, const size_t dims[3], int direction);

/* read_affine_translation (signal_model.rs:119-176): for each of n text files,
 * the first line containing "Parameters:" must hold 12 numbers (a 3x4 affine
 * transform); out (n x 3) receives its last three, times multiplier. */
llr_status llr_read_affine_translation(const char* const* files, size_t n, float multiplier, float* out);

/* ---------------------------------------------------------------- fftshift.rs */
/* fftshift / ifftshift (fftshift.rs:103-116): circular shift by floor(d/2) /
 * ceil(d/2) along each of the (up to 3) dims, generic element size. */
llr_status llr_fftshift(llr_ctx* ctx, void* data, size_t elem_bytes, const size_t dims[3]);
llr_status llr_ifftshift(llr_ctx* ctx, void* data, size_t elem_bytes, const size_t dims[3]);
/* phase_shift3 (fftshift.rs:42-74): x *= exp(i*sign*pi*sum_a (i_a - n_a/2)),
 * sign -1 for LLR_FORWARD, +1 for LLR_INVERSE, f32 phase as the reference. */
llr_status llr_phase_shift3(llr_ctx* ctx, llr_cf32* x, size_t len, const size_t dims[3], int direction);

/* ------------------------------------------------------------- lib.rs helpers */
/* wavelet_denoise (lib.rs:430-458) for n_volumes volumes: multi-level 3-D db2
 * analysis (levels < 0: floor(log2(min dim / 3))), soft threshold
 * shrink_complex (lib.rs:460-465) of every subband, synthesis. |x| = 0 maps to
 * 0 where the reference produces NaN. */
llr_status llr_wavelet_denoise(llr_ctx* ctx, llr_cf32* volumes, size_t n_volumes, const size_t vol_size[3],
                               float alpha, int levels);

/* wavelet_denoise of every volume moved circularly by shift (sample i of the
 * moved volume is sample i + shift, each axis modulo its length), then moved
 * back: shift-invariant (cycle-spinning) shrinkage, the use the reference's
 * shift hook (kernels.cu:80-95, 194, 264) was built for. */
llr_status llr_wavelet_denoise_shifted(llr_ctx* ctx, llr_cf32* volumes, size_t n_volumes, const size_t vol_size[3],
                                       float alpha, int levels, const int shift[3]);

/* cu_dwt3_axis (cudwt3.rs:250-298; kernels.cu:155-212, 270-310): one-level
 * analysis along `axis` of a col-major vol_size box into a decomp_size box
 * (decomp_size[axis] = 2 nc, nc = (n + F - 1) / 2; the other sizes equal):
 * symmetric extension, approximation at [0, nc) and detail at [nc, 2 nc).
 * Any even filter length F <= 32 (lo_d, hi_d on the host). shift (may be NULL)
 * moves every volume-side read circularly, each axis modulo its own length
 * (the reference uses the first dimension's length for all three,
 * kernels.cu:85-92). */
llr_status llr_dwt3_axis(llr_ctx* ctx, const llr_cf32* vol, size_t vol_len, llr_cf32* decomp, size_t decomp_len,
                         const size_t vol_size[3], const size_t decomp_size[3], int axis, const float* lo_d,
                         const float* hi_d, size_t filter_len, const int shift[3]);
/* cu_idwt3_axis (cudwt3.rs:300-351; kernels.cu:215-268, 312-348): the
 * synthesis, keeping the central 2 nc - F + 2 samples; shift moves the
 * volume-side writes. With equal shifts it inverts llr_dwt3_axis. */
llr_status llr_idwt3_axis(llr_ctx* ctx, llr_cf32* vol, size_t vol_len, const llr_cf32* decomp, size_t decomp_len,
                          const size_t vol_size[3], const size_t decomp_size[3], int axis, const float* lo_r,
                          const float* hi_r, size_t filter_len, const int shift[3]);
/* cu_dwt_denoise (cudwt3.rs:102-247): analysis along x, y, z then synthesis
 * along z, y, x of one volume, in place, each stage with the shift hook on its
 * input box; no shrinkage (a perfect-reconstruction round trip). */
llr_status llr_dwt_denoise(llr_ctx* ctx, llr_cf32* vol, const size_t vol_size[3], const float* lo_d,
                           const float* hi_d, const float* lo_r, const float* hi_r, size_t filter_len,
                           const int shift[3]);

/* hard_project (lib.rs:392-406): where y != 0, err += |x - y|^2 and x = y. */
llr_status llr_hard_project(llr_ctx* ctx, llr_cf32* x, const llr_cf32* y, size_t len, double* sum_sq_error);

/* calc_undersampling_factor (lib.rs:409-417). */
llr_status llr_calc_undersampling_factor(llr_ctx* ctx, const llr_cf32* y, size_t len, double* factor,
                                         double* total_energy);

/* under_sample_kspace (data_import.rs:416-438): zero each readout line whose
 * phase-encode mask entry (y, z, q) is zero; mask is ny x nz x nq complex. */
llr_status llr_under_sample_kspace(llr_ctx* ctx, llr_cf32* data, const size_t dims[3], size_t nq,
                                   const llr_cf32* mask);

/* ------------------------------------------------------- diffusion scalars */
/* The per-voxel part of calc_scalars (diffusion_models.rs:21-129): s0 = mean |x|
 * over the b0 volumes (|g| <= b0_tol, data_import.rs:322-328), log-linear
 * least-squares tensor over the other volumes (build_rhs :205-216, solve_tensor
 * :191-202, build_design_matrix :221-253), its eigenvalues (solve_eigenvalues
 * :160-186) and FA, MD, RD (:111-116). volumes: n_volumes x n_voxels complex;
 * b_vecs: n_volumes x 3. scalars: 6 planes of n_voxels floats, in the order
 * e1 >= e2 >= e3, FA, MD, RD (FA is NaN where all eigenvalues are 0, as in the
 * reference). The fit and eigen-solve run in FP64. */
llr_status llr_dti_scalars(llr_ctx* ctx, const llr_cf32* volumes, size_t n_voxels, size_t n_volumes,
                           const float* b_vals, const float* b_vecs, float b0_tol, float* scalars);
/* Same on device pointers (volumes, scalars); b_vals/b_vecs are host arrays.
 * Synchronous with the context's stream. */
llr_status llr_dti_scalars_device(llr_ctx* ctx, const llr_cf32* d_volumes, size_t n_voxels, size_t n_volumes,
                                  const float* b_vals, const float* b_vecs, float b0_tol, float* d_scalars);

/* ---------------------------------------------------- device-resident recon */
#define LLR_MODE_RANK 0 /* hard rank truncation (low_rank.rs:10-69) */
#define LLR_MODE_SVT 1  /* singular-value soft threshold */

#define LLR_PREC_FP32 0  /* plain form: FP32 throughout */
#define LLR_PREC_MIXED 1 /* BF16 MFMA Gram and projection (FP32 accumulation; with half storage
                          the projection runs on FP16 MFMA, the stored halves being its operand),
                          one FP32 Jacobi refinement sweep, FP64 eigen-solve for blocks with Gram
                          condition > 1e6 */

#define LLR_STORE_FP32 0 /* complex64 */
#define LLR_STORE_FP16 1 /* two IEEE half-precision numbers per complex value, FP32 arithmetic in
                            every kernel, data scaled by sqrt(voxels / measurement energy) while
                            stored (outputs, coefficients and thresholds in the caller's units);
                            LLR_PREC_MIXED with a subspace only (halves the HBM bytes per pass) */

typedef struct {
  size_t vol[3];         /* nx, ny, nz */
  size_t nq;             /* volumes */
  size_t blk[3];         /* llr_block_size (lib.rs:50) */
  size_t n_iter;         /* n_iter (lib.rs:62): maximum iterations */
  int llr_mode;          /* LLR_MODE_* */
  int precision;         /* LLR_PREC_* */
  size_t rank;           /* llr_block_rank (lib.rs:58) for LLR_MODE_RANK */
  double svt_lambda;     /* LLR_MODE_SVT: absolute threshold if > 0 ... */
  double lambda_factor;  /* ... else lambda_factor * sigma_max of the zero-filled blocks at shift 0 */
  double wavelet_alpha;  /* wavelet_shrink_factor (lib.rs:60); 0 disables the wavelet step */
  int wavelet_levels;    /* < 0: floor(log2(min dim / 3)) */
  size_t srank;          /* subspace rank; 0: no subspace (lib.rs:281-292): the LLR step then works
                            on all nq volumes, up to 8 in every form, up to 128 in the FP32 form */
  const llr_cf32* subspace; /* nq x srank, col-major (lib.rs:210-219); host pointer, copied */
  const size_t* shifts;  /* n_iter x 3 block shifts (lib.rs:266); host pointer, copied */
  double residual_stop;  /* > 0: stop once the relative residual falls to this value */
  int partial_blocks;    /* 1: ceil(vol/blk) block grid whose edge blocks are truncated at the
                            volume end (extension; the reference grid is floor(vol/blk),
                            block.rs:112-118, leaving the remainder voxels out of the LLR step) */
  int storage;           /* LLR_STORE_*: how the iterate, the measurements' working copy and the
                            wavelet workspace are held on the device (extension) */
  int no_hybrid;         /* 0: when every x-line of the measurements is fully measured or empty
                            (a phase-encode mask, as under_sample_kspace makes), the set-up takes
                            them through the inverse x transform once and the iterations skip the
                            x passes, which commute with the hard projection; 1: always the full
                            3D forward model and adjoint, as the reference (extension) */
  const float* lin_img_corr; /* nq x 3 linear image corrections of the signal model
                            (signal_model.rs:34-79: a sub-voxel image shift per volume, e.g. from
                            llr_read_affine_translation); host pointer, copied; NULL: zero, as
                            lib.rs:148 fixes them */
  int fused_projection;  /* where the subspace wavelet is fused into level 0 (a subspace, every
                            voxel in a block, the wavelet on) and the iteration is hybrid, the
                            adjoint's inverse y pass can hand the next LLR step the subspace
                            coefficients C = X S directly, the image X being formed once when
                            the run ends: 0 with complex64 storage (where it is faster), 1 with
                            either storage, -1 never (every iteration stores X and projects it,
                            as the reference) (extension) */
} llr_recon_params;

/* Create a reconstruction (allocates the iterate, the measurements and all
 * workspaces on the context's device). */
llr_status llr_recon_create(llr_ctx* ctx, const llr_recon_params* params, llr_recon** out);
void llr_recon_destroy(llr_recon* r);
/* Measurements (k-space, nx*ny*nz*nq complex64, zero = unsampled). on_device != 0:
 * y is a device pointer on the context's device. */
llr_status llr_recon_set_measurements(llr_recon* r, const llr_cf32* y, size_t len, int on_device);
/* Optional start of the iterate (image space) for the next run, otherwise the
 * zero-filled adjoint; the SVT threshold is then taken from this start. */
llr_status llr_recon_set_initial(llr_recon* r, const llr_cf32* x, size_t len, int on_device);
/* llr_recon_exec loop (lib.rs:227-341): estimate linear corrections, zero-filled
 * adjoint (unless a warm start was set), then n_iter iterations of
 * LLR -> wavelet -> forward model -> hard projection -> adjoint.
 * err_hist (may be NULL) receives err/total_energy per iteration (lib.rs:321);
 * iters_done (may be NULL) the number of iterations run. */
llr_status llr_recon_run(llr_recon* r, double* err_hist, size_t* iters_done);
/* Set-up of a run without iterating (lib.rs:224-258): total energy, k-space
 * corrections and ramp tables, the zero-filled adjoint (or the start set with
 * llr_recon_set_initial) and the SVT threshold. The next llr_recon_run starts
 * from this state instead of repeating it. */
llr_status llr_recon_prepare(llr_recon* r);
/* Subspace coefficients of the current iterate, C = X S (srank planes of
 * nx*ny*nz complex values, x fastest; lib.rs:285-286 over every voxel). Needs
 * a subspace. With llr_recon_prepare these are the zero-filled adjoint's. */
llr_status llr_recon_get_coefficients(llr_recon* r, llr_cf32* out, size_t len, int on_device);
/* Replace the iterate with C S^H for every voxel (srank planes). After
 * llr_recon_prepare the threshold of the zero-filled adjoint is kept. */
llr_status llr_recon_set_coefficients(llr_recon* r, const llr_cf32* c, size_t len, int on_device);

/* Copy the current image out (to a device pointer when on_device != 0). */
llr_status llr_recon_get_image(llr_recon* r, llr_cf32* out, size_t len, int on_device);
/* Threshold used by the SVT step and the per-volume k-space corrections found. */
llr_status llr_recon_info(llr_recon* r, double* lambda, float* ksp_corr);

/* 1 in *hybrid when the last set-up chose the (x, ky, kz) iteration (see
 * llr_recon_params.no_hybrid), else 0. No counterpart in the reference. */
llr_status llr_recon_hybrid(llr_recon* r, int* hybrid);

/* Mixed-precision statistics: LLR blocks whose K x K eigen-solve fell back to
 * FP64 (Gram condition number above 1e6) and blocks processed, since creation.
 * No counterpart in the reference (its SVD is FP32 only). */
llr_status llr_recon_mixed_stats(llr_recon* r, unsigned long long* fp64_blocks, unsigned long long* blocks);

/* Stage timing with HIP events on the context's stream (for roofline reports).
 * Stages: 0 LLR block kernel, 1 subspace projection, 2 wavelet (all its
 * kernels), 3 FFT x pass (forward), 4 FFT y pass (forward), 5 FFT z pass + hard
 * projection, 6 residual reduction, 7 run set-up (statistics, corrections,
 * zero-filled adjoint, threshold), 8 warm-start U-Net prediction (packing,
 * network, unpacking), 9 FFT y pass (adjoint), 10 FFT x pass (adjoint), and
 * the wavelet's kernels, also counted in stage 2: 11 the level-0 kernel that
 * reads the iterate or its subspace coefficients, 12 the level-0 kernels in
 * between, 13 the level-0 kernel that writes the iterate, 14 levels >= 1, and
 * 15 the adjoint's y pass fused with the subspace projection. */
#define LLR_NUM_STAGES 16
llr_status llr_recon_enable_timing(llr_recon* r, int enable);
llr_status llr_recon_stage_times(llr_recon* r, double* total_ms, size_t* launches);

/* ---------------------------------------------------- batched reconstructions */
/* Many reconstructions of uneven size (no counterpart in the reference, which
 * runs one llr_recon_exec at a time). Each member has its own parameters
 * (sizes, volumes, subspace, mode, threshold, shifts); members share n_iter,
 * run the FP32 form without residual_stop, and use partial edge blocks. One
 * iteration of every member is captured as a HIP graph: a projection per
 * member, ONE persistent LLR launch over the blocks of all members, then each
 * member's back-projection, wavelet, forward model, hard projection, adjoint
 * and residual on its own stream; the graph is replayed n_iter times.
 * autotune != 0: at the first run each member picks its block size from
 * {6,8,10}^3, the LLR kernel's LDS staging and the FFT tiles per workgroup by
 * timing them on its own data. */
typedef struct llr_recon_batch llr_recon_batch;
llr_status llr_recon_batch_create(llr_ctx* ctx, const llr_recon_params* params, size_t n_recons, int autotune,
                                  llr_recon_batch** out);
void llr_recon_batch_destroy(llr_recon_batch* batch);
llr_status llr_recon_batch_set_measurements(llr_recon_batch* batch, size_t index, const llr_cf32* y, size_t len,
                                            int on_device);
/* err_hist: n_recons x n_iter relative residuals (or null). */
llr_status llr_recon_batch_run(llr_recon_batch* batch, double* err_hist);
llr_status llr_recon_batch_get_image(llr_recon_batch* batch, size_t index, llr_cf32* out, size_t len, int on_device);
/* What the autotuner chose for a member (and its SVT threshold). */
llr_status llr_recon_batch_tuning(llr_recon_batch* batch, size_t index, size_t blk[3], int fft_tpw[3],
                                  int* llr_stage, double* lambda);

/* --------------------------------------------------------- learned warm start */
/* A 3D U-Net in subspace-coefficient space (no counterpart in the reference).
 * levels L, widths base*2^l; per level two 3^3 conv + ReLU, 2x2x2 average
 * pooling, nearest upsampling with [upsampled, skip] concatenation, a 1^3
 * output convolution added to the input. BF16 weights and activations, MFMA
 * with FP32 accumulation, hand-written forward and backward kernels, Adam
 * (beta1 0.9, beta2 0.999, eps 1e-8) on FP32 master weights.
 * Tensors are (nz, ny, nx, channels) FP32 with channels fastest; sizes must be
 * multiples of 2^(L-1). */
typedef struct llr_unet llr_unet;
llr_status llr_unet_create(llr_ctx* ctx, int in_channels, int base_channels, int levels, unsigned long long seed,
                           llr_unet** out);
void llr_unet_destroy(llr_unet* net);
llr_status llr_unet_num_params(llr_unet* net, size_t* n);
/* Parameters in layer order (enc0a, enc0b, ..., dec{L-2}a, dec{L-2}b, ..., dec0b, out), each
 * w[o][co][ci] (o = (dz*k + dy)*k + dx) then b[co]. Setting them resets Adam. */
llr_status llr_unet_get_params(llr_unet* net, float* params, size_t n);
llr_status llr_unet_set_params(llr_unet* net, const float* params, size_t n);
/* dims = {nx, ny, nz}. x, y: nx*ny*nz*in_channels floats (host or device). */
llr_status llr_unet_forward(llr_unet* net, const float* x, float* y, const size_t dims[3], int on_device);
/* Gradients of sum(gy * forward(x)) w.r.t. the parameters (host arrays). */
llr_status llr_unet_gradients(llr_unet* net, const float* x, const float* gy, const size_t dims[3], float* grads);
/* One Adam step on the mean squared error between forward(x) and target. */
llr_status llr_unet_train_step(llr_unet* net, const float* x, const float* target, const size_t dims[3], float lr,
                               int on_device, double* loss);
/* Copy of a stored BF16 activation of the last forward pass (as floats), for
 * inspection and tests: kind 0/1 encoder conv a/b outputs, 2 pooled input of
 * the level, 3 decoder concatenation, 4/5 decoder conv a/b outputs; n must be
 * the tensor's element count. */
llr_status llr_unet_get_activation(llr_unet* net, int kind, int level, float* out, size_t n);
/* Coefficient <-> channel mapping (K <= 8): per voxel the network sees
 * b = to_net * c / scale as channels [Re b0, Im b0, Re b1, ...] (remaining
 * channels zero), and a prediction p maps back to from_net * p * scale.
 * to_net / from_net are K x K real, row-major host arrays; NULL = identity. */
/* Device planes of K complex coefficients (x fastest) -> predicted planes (may alias). */
llr_status llr_unet_predict_coefficients(llr_unet* net, const llr_cf32* d_coef, size_t K, const size_t vol[3],
                                         float scale, const float* to_net, const float* from_net, llr_cf32* d_out);
/* K complex planes -> (voxels, channels) network tensor on the device, FP32
 * (bf16_out = 0) or BF16 (bf16_out = 1); channels a multiple of 16. */
llr_status llr_unet_pack_coefficients(llr_ctx* ctx, const llr_cf32* d_coef, size_t K, const size_t vol[3],
                                      int channels, float scale, const float* to_net, int bf16_out, void* d_out);
/* The warm start's input scale: sqrt(mean over voxels of |(to_net * c)_0|^2). */
llr_status llr_unet_coefficient_scale(llr_ctx* ctx, const llr_cf32* d_coef, size_t K, const size_t vol[3],
                                      const float* to_net, double* scale);
/* One Adam step on a crop of stored BF16 (input, target) tensors (device,
 * (nz, ny, nx, in_channels) each): crop_dims voxels at origin ({x, y, z}). */
llr_status llr_unet_train_step_crop(llr_unet* net, const void* x_full, const void* t_full, const size_t full_dims[3],
                                    const size_t origin[3], const size_t crop_dims[3], float lr, double* loss);

/* Learned warm start of a reconstruction with a subspace. The zero-filled
 * adjoint is computed as for a cold start (its SVT threshold is kept), its
 * subspace coefficients go through the network (mapping above, scale from
 * llr_unet_coefficient_scale), and the iterate becomes the prediction's
 * back-projection made data-consistent (forward model, hard projection,
 * adjoint), as every iterate that enters the LLR step is. llr_recon_run then
 * continues from it. scale and residual (may be NULL) receive the input scale
 * and the prediction's relative data-consistency error. */
llr_status llr_recon_warm_start(llr_recon* r, llr_unet* net, const float* to_net, const float* from_net,
                                double* scale, double* residual);
typedef struct {
  const float* to_net;   /* K x K, see above (NULL = identity) */
  const float* from_net; /* K x K (NULL = identity) */
  double residual_stop;  /* the iterations stop once the relative residual falls to this value */
  size_t max_iter;       /* iteration budget of the warm-started run (0 or > n_iter: n_iter) */
  double momentum;       /* b in [0, 1): the LLR step of the warm-started run acts on
                            x_k + b (x_k - x_{k-1}) (needs the fused wavelet: a subspace and a
                            block grid covering every voxel); 0: plain iterations */
  size_t min_iter;       /* the residual_stop test starts after this many iterations */
} llr_warm_params;
/* Warm start, then iterations (at least min_iter) until residual_stop. If the
 * warm-started run does not reach residual_stop within max_iter iterations the
 * cold start is run instead (zero-filled adjoint, n_iter plain iterations) and
 * kept: fell_back = 1. err_hist (n_iter entries) receives the kept run's
 * relative residuals. */
llr_status llr_recon_run_learned(llr_recon* r, llr_unet* net, const llr_warm_params* wp, double* err_hist,
                                 size_t* iters_done, int* fell_back);

/* ------------------------------------------------- data ingest and egress
 * Host-only (no device work): what llr_recon_exec reads and writes around the
 * reconstruction. Paths may be given with or without their extension. */

/* BART .cfl/.hdr (the cfl crate's get_dims / read_to_buffer / write_buffer,
 * data_import.rs:37-185, lib.rs:354-362): the .hdr holds "# Dimensions" and up
 * to 16 sizes, the .cfl the column-major complex64 samples, little endian.
 * llr_cfl_dims returns the sizes written (dims[i] = 1 beyond *ndims); len must
 * equal their product. */
llr_status llr_cfl_dims(const char* path, size_t dims[16], size_t* ndims);
llr_status llr_cfl_read(const char* path, llr_cf32* data, size_t len);
llr_status llr_cfl_write(const char* path, const size_t* dims, size_t ndims, const llr_cf32* data, size_t len);

/* DtiDataSetBuilder::with_b_table's parser (data_import.rs:110-145): the .txt
 * file holds whitespace-separated [b, gx, gy, gz] rows; *n receives the row
 * count (b_vals and b_vecs both null: size query). b_vecs is n x 3. */
llr_status llr_read_b_table(const char* path, float* b_vals, float* b_vecs, size_t capacity, size_t* n);
/* b0_mask (data_import.rs:322-328): mask[i] = |g_i| <= tolerance (tolerance >= 0). */
llr_status llr_b0_mask(const float* b_vecs, size_t n, float tolerance, int* mask);

/* DtiDataSetBuilder (data_import.rs:13-273) over a directory of per-volume
 * .cfl files: the files matching <dir>/<pattern>.cfl (glob), sorted
 * ascending, every one with the same 3 non-singleton dimensions; expected = 0
 * accepts any count (from_cfl_volume_dir, :67-108). */
#define LLR_VOLUMES_DTI 0 /* b-table rows with |g| > tolerance (load_dti_set, dti_indices) */
#define LLR_VOLUMES_B0 1  /* b0 rows (load_b0_set, b0_indices) */
#define LLR_VOLUMES_ALL 2 /* every file (load_all) */
typedef struct llr_dataset llr_dataset;
llr_status llr_dataset_from_cfl_dir(const char* dir, const char* pattern, size_t expected, llr_dataset** out);
void llr_dataset_destroy(llr_dataset* ds);
/* with_b_table (:110-145): one row per file, b0 volumes by b0_tol. */
llr_status llr_dataset_with_b_table(llr_dataset* ds, const char* path, float b0_tol);
/* with_sample_mask (:187-198): an (ny, nz, n_files) .cfl mask. */
llr_status llr_dataset_with_sample_mask(llr_dataset* ds, const char* path);
llr_status llr_dataset_info(const llr_dataset* ds, size_t vol[3], size_t* n_volumes, size_t* n_dti, size_t* n_b0);
/* name of file `index` (NUL-terminated, cap bytes) */
llr_status llr_dataset_file(const llr_dataset* ds, size_t index, char* out, size_t cap);
/* file indices of a set (dti_indices / b0_indices, :253-265); out null: count only */
llr_status llr_dataset_indices(const llr_dataset* ds, int which, size_t* out, size_t cap, size_t* n);
/* b-values (n_files) and vectors (n_files x 3) of every file */
llr_status llr_dataset_b_table(const llr_dataset* ds, float* b_vals, float* b_vecs);
/* the selected volumes, contiguous (load_volumes, :173-185); len = nx*ny*nz*count */
llr_status llr_dataset_load(const llr_dataset* ds, int which, llr_cf32* out, size_t len);
/* the selected volumes' masks (dti_sample_mask / b0_sample_mask, :204-250): ny*nz*count */
llr_status llr_dataset_sample_mask(const llr_dataset* ds, int which, llr_cf32* out, size_t len);

/* LlrReconParams (lib.rs:47-80) as its TOML file holds it, plus keys this
 * library adds (all optional): llr_mode = "rank" | "svt", svt_lambda_factor,
 * precision = "fp32" | "mixed", storage = "fp32" | "fp16", shift_seed. */
typedef struct {
  size_t llr_block_size[3];
  int has_max_rand_shift;
  size_t max_rand_shift[3];
  size_t fft_max_batch_size;     /* accepted; the resident reconstruction does not batch */
  size_t cas_mat_max_batch_size; /* accepted; all blocks of an iteration run in one launch */
  size_t llr_block_rank;
  float wavelet_shrink_factor;
  size_t n_iter;
  int has_subspace;
  char subspace_projection_mat[4096];
  int llr_mode;            /* LLR_MODE_RANK (the reference's) or LLR_MODE_SVT */
  double lambda_factor;    /* SVT threshold factor (of sigma_max of the zero-filled blocks) */
  int precision;           /* LLR_PREC_* */
  int storage;             /* LLR_STORE_* */
  unsigned long long seed; /* block-shift generator seed (the reference's shifts are unseeded) */
} llr_toml_params;
/* the defaults of lib.rs:67-80 (and of the added keys: rank mode, 0.02, FP32, seed 0) */
void llr_recon_params_default(llr_toml_params* p);
/* read <path>.toml (bin/llr-recon.rs:8-12): the reference's required fields must be
 * there; unknown keys are ignored */
llr_status llr_recon_params_read(const char* path, llr_toml_params* p);
/* write <path>.toml with the reference's fields (bin/llr-reco-params.rs:14-23) */
llr_status llr_recon_params_write(const char* path, const llr_toml_params* p);
/* error_history.json (lib.rs:345-351): the relative residuals as a pretty JSON array */
llr_status llr_write_error_history(const char* path, const double* hist, size_t n);

#ifdef __cplusplus
}
#endif

#endif /* LLR_HIP_H */
