// Diffusion-tensor scalars on gfx950: the per-voxel part of calc_scalars
// (diffusion_models.rs:21-129).
//
// Per voxel: s0 = mean |x| over the b0 volumes (:42-50), clamped to FLT_MIN;
// rhs_j = ln(clamp(|x_j| / s0, FLT_MIN, 1)) over the diffusion-weighted volumes
// (build_rhs, :205-216); the six tensor entries {Dxx, Dyy, Dzz, Dxy, Dxz, Dyz}
// are the least-squares solution of A d = rhs (solve_tensor, :191-202) with the
// design matrix of build_design_matrix (:221-253). A is the same for every
// voxel, so its pseudo-inverse P = (A^T A)^-1 A^T is formed once on the host
// (long double) and each voxel's fit is a 6 x n_dw product. The eigenvalues of
// the symmetric tensor (solve_eigenvalues, :160-186) come from the closed-form
// trigonometric solution of the characteristic cubic (O. K. Smith, "Eigenvalues
// of a symmetric 3x3 matrix", CACM 4(4), 1961). Outputs e1 >= e2 >= e3, FA, MD
// and RD (:111-116). Everything past the magnitudes runs in FP64.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <vector>

#include "common.h"
#include "internal.h"

namespace {

constexpr int DTI_MAX_DW = 256;

__device__ inline void sym3_eigvals(const double* d, double* e) {
  // d = {Dxx, Dyy, Dzz, Dxy, Dxz, Dyz}
  const double a00 = d[0], a11 = d[1], a22 = d[2], a01 = d[3], a02 = d[4], a12 = d[5];
  const double p1 = a01 * a01 + a02 * a02 + a12 * a12;
  const double q = (a00 + a11 + a22) / 3.0;
  if (p1 == 0.0) {
    double x = a00, y = a11, z = a22, t;
    if (x < y) { t = x; x = y; y = t; }
    if (y < z) { t = y; y = z; z = t; }
    if (x < y) { t = x; x = y; y = t; }
    e[0] = x; e[1] = y; e[2] = z;
    return;
  }
  const double b00 = a00 - q, b11 = a11 - q, b22 = a22 - q;
  const double p2 = b00 * b00 + b11 * b11 + b22 * b22 + 2.0 * p1;
  const double p = sqrt(p2 / 6.0);
  // r = det((A - qI) / p) / 2
  const double det = b00 * (b11 * b22 - a12 * a12) - a01 * (a01 * b22 - a12 * a02) + a02 * (a01 * a12 - b11 * a02);
  double r = det / (2.0 * p * p * p);
  r = fmin(fmax(r, -1.0), 1.0);
  const double phi = acos(r) / 3.0;
  const double two_pi_3 = 2.0943951023931954923;
  const double e1 = q + 2.0 * p * cos(phi);
  const double e3 = q + 2.0 * p * cos(phi + two_pi_3);
  e[0] = e1;
  e[2] = e3;
  e[1] = 3.0 * q - e1 - e3;
}

__global__ void __launch_bounds__(256) k_dti_scalars(const float2* __restrict__ vols, long long N,
                                                     const int* __restrict__ b0_idx, int nb0,
                                                     const int* __restrict__ dw_idx, int ndw,
                                                     const double* __restrict__ P, float* __restrict__ out) {
  __shared__ double sP[6 * DTI_MAX_DW];
  __shared__ int sdw[DTI_MAX_DW];
  for (int i = threadIdx.x; i < 6 * ndw; i += blockDim.x) sP[i] = P[i];
  for (int i = threadIdx.x; i < ndw; i += blockDim.x) sdw[i] = dw_idx[i];
  __syncthreads();
  const double tiny = (double)FLT_MIN;
  for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < N; v += (long long)gridDim.x * blockDim.x) {
    double s0 = 0.0;
    for (int i = 0; i < nb0; ++i) {
      float2 x = vols[(long long)b0_idx[i] * N + v];
      s0 += sqrt((double)x.x * x.x + (double)x.y * x.y);
    }
    s0 = fmax(s0 / (double)nb0, tiny);
    double d[6] = {0, 0, 0, 0, 0, 0};
    for (int j = 0; j < ndw; ++j) {
      float2 x = vols[(long long)sdw[j] * N + v];
      double a = sqrt((double)x.x * x.x + (double)x.y * x.y);
      double r = log(fmin(fmax(a / s0, tiny), 1.0));
#pragma unroll
      for (int k = 0; k < 6; ++k) d[k] += sP[k * ndw + j] * r;
    }
    double e[3];
    sym3_eigvals(d, e);
    const double md = (e[0] + e[1] + e[2]) / 3.0;
    const double rd = (e[1] + e[2]) / 2.0;
    const double num = sqrt((e[0] - md) * (e[0] - md) + (e[1] - md) * (e[1] - md) + (e[2] - md) * (e[2] - md));
    const double den = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    const double fa = sqrt(1.5) * num / den;  // 0/0 -> NaN, as the reference
    out[v] = (float)e[0];
    out[N + v] = (float)e[1];
    out[2 * N + v] = (float)e[2];
    out[3 * N + v] = (float)fa;
    out[4 * N + v] = (float)md;
    out[5 * N + v] = (float)rd;
  }
}

}  // namespace

namespace llr {

// Pseudo-inverse (A^T A)^-1 A^T of the n_dw x 6 design matrix, row-major 6 x n_dw.
static llr_status dti_pinv(const std::vector<long double>& A, int m, std::vector<double>& P) {
  long double G[36] = {0}, L[36] = {0};
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j)
      for (int r = 0; r < m; ++r) G[i * 6 + j] += A[r * 6 + i] * A[r * 6 + j];
  // Cholesky G = L L^T
  long double gmax = 0;
  for (int i = 0; i < 6; ++i) gmax = std::max(gmax, G[i * 6 + i]);
  for (int j = 0; j < 6; ++j) {
    long double s = G[j * 6 + j];
    for (int k = 0; k < j; ++k) s -= L[j * 6 + k] * L[j * 6 + k];
    if (!(s > 1e-12L * gmax)) return fail(LLR_ERR_ARG, "the b-table does not determine the diffusion tensor");
    L[j * 6 + j] = std::sqrt(s);
    for (int i = j + 1; i < 6; ++i) {
      long double t = G[i * 6 + j];
      for (int k = 0; k < j; ++k) t -= L[i * 6 + k] * L[j * 6 + k];
      L[i * 6 + j] = t / L[j * 6 + j];
    }
  }
  P.assign((size_t)6 * m, 0.0);
  for (int r = 0; r < m; ++r) {
    long double y[6], z[6];
    for (int i = 0; i < 6; ++i) {  // L y = a_r
      long double t = A[r * 6 + i];
      for (int k = 0; k < i; ++k) t -= L[i * 6 + k] * y[k];
      y[i] = t / L[i * 6 + i];
    }
    for (int i = 5; i >= 0; --i) {  // L^T z = y
      long double t = y[i];
      for (int k = i + 1; k < 6; ++k) t -= L[k * 6 + i] * z[k];
      z[i] = t / L[i * 6 + i];
    }
    for (int i = 0; i < 6; ++i) P[(size_t)i * m + r] = (double)z[i];
  }
  return LLR_OK;
}

llr_status dti_scalars_device(llr_ctx* ctx, const float2* vols, size_t n_vox, size_t nq, const float* bvals,
                              const float* bvecs, float b0_tol, float* out) {
  LLR_CHECK(b0_tol >= 0.f, "vector magnitude tolerance must be non-negative");
  std::vector<int> b0, dw;
  for (size_t q = 0; q < nq; ++q) {
    const float* g = bvecs + 3 * q;
    float nrm = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);  // data_import.rs:322-328
    (nrm <= b0_tol ? b0 : dw).push_back((int)q);
  }
  LLR_CHECK(!b0.empty(), "the data set has no b0 volume");
  if ((int)dw.size() > DTI_MAX_DW) return fail(LLR_ERR_UNSUPPORTED, "more than 256 diffusion-weighted volumes");
  const int m = (int)dw.size();
  std::vector<long double> A((size_t)m * 6);
  for (int r = 0; r < m; ++r) {  // diffusion_models.rs:233-252
    const long double b = bvals[dw[r]];
    const long double gx = bvecs[3 * dw[r]], gy = bvecs[3 * dw[r] + 1], gz = bvecs[3 * dw[r] + 2];
    A[r * 6 + 0] = -b * gx * gx;
    A[r * 6 + 1] = -b * gy * gy;
    A[r * 6 + 2] = -b * gz * gz;
    A[r * 6 + 3] = -2 * b * gx * gy;
    A[r * 6 + 4] = -2 * b * gx * gz;
    A[r * 6 + 5] = -2 * b * gy * gz;
  }
  std::vector<double> P;
  llr_status st = dti_pinv(A, m, P);
  if (st) return st;
  DevBuf dP, db0, ddw;
  LLR_HIP(dP.alloc(P.size() * sizeof(double)));
  LLR_HIP(db0.alloc(b0.size() * sizeof(int)));
  LLR_HIP(ddw.alloc(dw.size() * sizeof(int)));
  LLR_HIP(hipMemcpyAsync(dP.p, P.data(), P.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  LLR_HIP(hipMemcpyAsync(db0.p, b0.data(), b0.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  LLR_HIP(hipMemcpyAsync(ddw.p, dw.data(), dw.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_dti_scalars, dim3(grid_for(n_vox, 256, 16384u)), dim3(256), 0, ctx->stream, vols,
                     (long long)n_vox, db0.as<int>(), (int)b0.size(), ddw.as<int>(), m, dP.as<double>(), out);
  LLR_HIP(hipGetLastError());
  // the small tables are freed when this call returns: wait for the kernel
  LLR_HIP(hipStreamSynchronize(ctx->stream));
  return LLR_OK;
}

}  // namespace llr
