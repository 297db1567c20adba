// Single-level wavelet transform along one axis with any filter pair, and the
// per-axis building blocks of the reference's GPU wavelet round trip:
//   dwt3_axis / idwt3_axis (kernels.cu:155-268, 270-349; cudwt3.rs:250-351),
//   cu_dwt_denoise (cudwt3.rs:102-247: analysis along x, y, z then synthesis
//   along z, y, x, no shrinkage),
// plus a cycle-spinning multi-level wavelet shrinkage (the shift hook of
// kernels.cu:80-95 used as intended).
//
// Conventions (the reference's): symmetric (half-sample) extension,
// nc = (n + F - 1) / 2 coefficients per band, approximation at [0, nc) and
// detail at [nc, 2 nc) of the transformed axis; the analysis reads sample
// sym(2c - (F - 1) + j + 1) with tap F - 1 - j; the synthesis keeps the
// central 2 nc - F + 2 samples of the full-length reconstruction.
//
// Shift hook: every sample address on the volume side is moved circularly by
// `shift` before it is read (analysis) or written (synthesis), each axis
// modulo its own length. The reference takes the first dimension's length as
// the modulus of all three axes (kernels.cu:85-92), which is wrong for
// non-cubic boxes; this is the intended per-axis form. With the same shift
// an analysis followed by the synthesis is the identity.
#include <string>

#include "common.h"
#include "internal.h"

using namespace llr;

namespace {

constexpr int kMaxTaps = 32;

struct DwtArgs {
  const float2* in;
  float2* out;
  long long vd[3];  // volume-side box (the signal)
  long long cd[3];  // coefficient-side box (2 nc along the axis)
  int axis, F, nc;
  int shift[3];
  float lo[kMaxTaps], hi[kMaxTaps];
};

__device__ __forceinline__ long long wrap(long long i, long long n) {
  long long r = i % n;
  return r < 0 ? r + n : r;
}

// volume-side address of (x, y, z) moved by the shift hook
__device__ __forceinline__ long long vaddr(const DwtArgs& a, long long x, long long y, long long z) {
  x = wrap(x + a.shift[0], a.vd[0]);
  y = wrap(y + a.shift[1], a.vd[1]);
  z = wrap(z + a.shift[2], a.vd[2]);
  return x + a.vd[0] * (y + a.vd[1] * z);
}

__device__ __forceinline__ int sym_index(int i, int n) {
  while (i < 0 || i >= n) i = i < 0 ? -i - 1 : 2 * n - i - 1;
  return i;
}

// one thread per (lane, coefficient c) of the coefficient box: writes the
// approximation and detail coefficients c of its lane
__global__ void __launch_bounds__(256) k_dwt_axis(DwtArgs a) {
  const long long lanes = a.cd[0] * a.cd[1] * a.cd[2] / a.cd[a.axis];
  const long long total = lanes * a.nc;
  const int n = (int)a.vd[a.axis];
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long lane = e / a.nc;
    const int c = (int)(e - lane * a.nc);
    // lane -> the two other coordinates (in axis order)
    long long p[3];
    const int a1 = a.axis == 0 ? 1 : 0, a2 = a.axis == 2 ? 1 : 2;
    p[a1] = lane % a.cd[a1];
    p[a2] = lane / a.cd[a1];
    float2 lo = make_float2(0.f, 0.f), hi = make_float2(0.f, 0.f);
    for (int j = 0; j < a.F; ++j) {
      p[a.axis] = sym_index(2 * c - (a.F - 1) + j + 1, n);
      const float2 v = a.in[vaddr(a, p[0], p[1], p[2])];
      const float fl = a.lo[a.F - 1 - j], fh = a.hi[a.F - 1 - j];
      lo.x += fl * v.x;
      lo.y += fl * v.y;
      hi.x += fh * v.x;
      hi.y += fh * v.y;
    }
    p[a.axis] = c;
    const long long s = (a.axis == 0 ? 1 : a.axis == 1 ? a.cd[0] : a.cd[0] * a.cd[1]);
    const long long base = p[0] + a.cd[0] * (p[1] + a.cd[1] * p[2]);
    a.out[base] = lo;
    a.out[base + (long long)a.nc * s] = hi;
  }
}

// one thread per volume-side sample (lane, s): the kept central sample s of
// the reconstruction from the approximation and detail bands of its lane
__global__ void __launch_bounds__(256) k_idwt_axis(DwtArgs a) {
  const long long total = a.vd[0] * a.vd[1] * a.vd[2];
  const int n = (int)a.vd[a.axis];
  const int full = 2 * a.nc + a.F - 1, keep = 2 * a.nc - a.F + 2, start = (full - keep) / 2;
  const long long cs = (a.axis == 0 ? 1 : a.axis == 1 ? a.cd[0] : a.cd[0] * a.cd[1]);
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    long long p[3] = {e % a.vd[0], (e / a.vd[0]) % a.vd[1], e / (a.vd[0] * a.vd[1])};
    const int smp = (int)p[a.axis];
    if (smp >= n) continue;
    float2 r = make_float2(0.f, 0.f);
    long long q[3] = {p[0], p[1], p[2]};
    for (int j = 0; j < a.F; ++j) {
      const int idx = start + smp - j;
      if (idx < 0 || idx >= 2 * a.nc || (idx & 1)) continue;
      q[a.axis] = idx / 2;
      const long long b = q[0] + a.cd[0] * (q[1] + a.cd[1] * q[2]);
      const float2 av = a.in[b], dv = a.in[b + (long long)a.nc * cs];
      r.x += a.lo[j] * av.x + a.hi[j] * dv.x;
      r.y += a.lo[j] * av.y + a.hi[j] * dv.y;
    }
    a.out[vaddr(a, p[0], p[1], p[2])] = r;
  }
}

llr_status make_args(DwtArgs& a, const size_t* vol, const size_t* dec, int axis, const float* f0, const float* f1,
                     size_t F, const int* shift) {
  if (axis < 0 || axis > 2) return fail(LLR_ERR_ARG, "axis must be 0, 1 or 2");
  if (F < 2 || F > (size_t)kMaxTaps || F % 2) return fail(LLR_ERR_UNSUPPORTED, "filter length must be even, 2..32");
  if (!f0 || !f1) return fail(LLR_ERR_ARG, "null filter");
  for (int d = 0; d < 3; ++d) {
    if (vol[d] == 0) return fail(LLR_ERR_ARG, "empty volume");
    const size_t want = d == axis ? 2 * ((vol[d] + F - 1) / 2) : vol[d];
    if (dec[d] != want)
      return fail(LLR_ERR_ARG, "decomposition size along axis " + std::to_string(d) + " must be " +
                                   std::to_string(want));
    a.vd[d] = (long long)vol[d];
    a.cd[d] = (long long)dec[d];
    a.shift[d] = shift ? shift[d] : 0;
  }
  if ((double)dec[0] * dec[1] * dec[2] >= 9.0e18) return fail(LLR_ERR_UNSUPPORTED, "box too large");
  a.axis = axis;
  a.F = (int)F;
  a.nc = (int)(dec[axis] / 2);
  for (size_t k = 0; k < F; ++k) {
    a.lo[k] = f0[k];
    a.hi[k] = f1[k];
  }
  return LLR_OK;
}

llr_status dwt_axis_device(llr_ctx* ctx, const float2* vol, float2* dec, const size_t* vs, const size_t* ds, int axis,
                           const float* lo, const float* hi, size_t F, const int* shift) {
  DwtArgs a{};
  llr_status st = make_args(a, vs, ds, axis, lo, hi, F, shift);
  if (st) return st;
  a.in = vol;
  a.out = dec;
  const size_t total = ds[0] * ds[1] * ds[2] / 2;
  hipLaunchKernelGGL(k_dwt_axis, dim3(grid_for(total, 256)), dim3(256), 0, ctx->stream, a);
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

llr_status idwt_axis_device(llr_ctx* ctx, float2* vol, const float2* dec, const size_t* vs, const size_t* ds, int axis,
                            const float* lo, const float* hi, size_t F, const int* shift) {
  DwtArgs a{};
  llr_status st = make_args(a, vs, ds, axis, lo, hi, F, shift);
  if (st) return st;
  a.in = dec;
  a.out = vol;
  const size_t total = vs[0] * vs[1] * vs[2];
  hipLaunchKernelGGL(k_idwt_axis, dim3(grid_for(total, 256)), dim3(256), 0, ctx->stream, a);
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

llr_status to_device(DevBuf& b, const void* h, size_t bytes) {
  LLR_HIP(b.alloc(bytes));
  if (h && bytes) LLR_HIP(hipMemcpy(b.p, h, bytes, hipMemcpyHostToDevice));
  return LLR_OK;
}

llr_status to_host(void* h, const DevBuf& b, size_t bytes, llr_ctx* ctx) {
  LLR_HIP(hipStreamSynchronize(ctx->stream));
  if (bytes) LLR_HIP(hipMemcpy(h, b.p, bytes, hipMemcpyDeviceToHost));
  return LLR_OK;
}

size_t prod3(const size_t* d) { return d[0] * d[1] * d[2]; }

}  // namespace

extern "C" {

llr_status llr_dwt3_axis(llr_ctx* ctx, const llr_cf32* vol, size_t vol_len, llr_cf32* decomp, size_t decomp_len,
                         const size_t vol_size[3], const size_t decomp_size[3], int axis, const float* lo_d,
                         const float* hi_d, size_t filter_len, const int shift[3]) {
  if (!ctx || !vol || !decomp || !vol_size || !decomp_size) return fail(LLR_ERR_ARG, "null argument");
  if (vol_len != prod3(vol_size)) return fail(LLR_ERR_ARG, "volume length does not match its size");
  if (decomp_len != prod3(decomp_size)) return fail(LLR_ERR_ARG, "decomposition length does not match its size");
  DwtArgs probe{};
  llr_status st = make_args(probe, vol_size, decomp_size, axis, lo_d, hi_d, filter_len, shift);
  if (st) return st;
  LLR_HIP(hipSetDevice(ctx->device));
  DevBuf dv, dd;
  if ((st = to_device(dv, vol, vol_len * 8))) return st;
  if ((st = to_device(dd, nullptr, decomp_len * 8))) return st;
  if ((st = dwt_axis_device(ctx, dv.as<float2>(), dd.as<float2>(), vol_size, decomp_size, axis, lo_d, hi_d,
                            filter_len, shift)))
    return st;
  return to_host(decomp, dd, decomp_len * 8, ctx);
}

llr_status llr_idwt3_axis(llr_ctx* ctx, llr_cf32* vol, size_t vol_len, const llr_cf32* decomp, size_t decomp_len,
                          const size_t vol_size[3], const size_t decomp_size[3], int axis, const float* lo_r,
                          const float* hi_r, size_t filter_len, const int shift[3]) {
  if (!ctx || !vol || !decomp || !vol_size || !decomp_size) return fail(LLR_ERR_ARG, "null argument");
  if (vol_len != prod3(vol_size)) return fail(LLR_ERR_ARG, "volume length does not match its size");
  if (decomp_len != prod3(decomp_size)) return fail(LLR_ERR_ARG, "decomposition length does not match its size");
  DwtArgs probe{};
  llr_status st = make_args(probe, vol_size, decomp_size, axis, lo_r, hi_r, filter_len, shift);
  if (st) return st;
  LLR_HIP(hipSetDevice(ctx->device));
  DevBuf dv, dd;
  if ((st = to_device(dd, decomp, decomp_len * 8))) return st;
  if ((st = to_device(dv, nullptr, vol_len * 8))) return st;
  if ((st = idwt_axis_device(ctx, dv.as<float2>(), dd.as<float2>(), vol_size, decomp_size, axis, lo_r, hi_r,
                             filter_len, shift)))
    return st;
  return to_host(vol, dv, vol_len * 8, ctx);
}

// cu_dwt_denoise (cudwt3.rs:102-247): x -> tmp1 -> tmp2 -> decomp along x, y, z,
// then back along z, y, x, each stage with the shift hook on its input box
llr_status llr_dwt_denoise(llr_ctx* ctx, llr_cf32* vol, const size_t vol_size[3], const float* lo_d,
                           const float* hi_d, const float* lo_r, const float* hi_r, size_t filter_len,
                           const int shift[3]) {
  if (!ctx || !vol || !vol_size) return fail(LLR_ERR_ARG, "null argument");
  if (filter_len < 2 || filter_len % 2 || filter_len > (size_t)kMaxTaps)
    return fail(LLR_ERR_UNSUPPORTED, "filter length must be even, 2..32");
  size_t t1[3] = {2 * ((vol_size[0] + filter_len - 1) / 2), vol_size[1], vol_size[2]};
  size_t t2[3] = {t1[0], 2 * ((vol_size[1] + filter_len - 1) / 2), vol_size[2]};
  size_t dc[3] = {t1[0], t2[1], 2 * ((vol_size[2] + filter_len - 1) / 2)};
  LLR_HIP(hipSetDevice(ctx->device));
  DevBuf dv, d1, d2, dd;
  llr_status st;
  const size_t n = prod3(vol_size);
  if ((st = to_device(dv, vol, n * 8))) return st;
  if ((st = to_device(d1, nullptr, prod3(t1) * 8))) return st;
  if ((st = to_device(d2, nullptr, prod3(t2) * 8))) return st;
  if ((st = to_device(dd, nullptr, prod3(dc) * 8))) return st;
  float2 *v = dv.as<float2>(), *a = d1.as<float2>(), *b = d2.as<float2>(), *c = dd.as<float2>();
  const size_t F = filter_len;
  if ((st = dwt_axis_device(ctx, v, a, vol_size, t1, 0, lo_d, hi_d, F, shift))) return st;
  if ((st = dwt_axis_device(ctx, a, b, t1, t2, 1, lo_d, hi_d, F, shift))) return st;
  if ((st = dwt_axis_device(ctx, b, c, t2, dc, 2, lo_d, hi_d, F, shift))) return st;
  if ((st = idwt_axis_device(ctx, b, c, t2, dc, 2, lo_r, hi_r, F, shift))) return st;
  if ((st = idwt_axis_device(ctx, a, b, t1, t2, 1, lo_r, hi_r, F, shift))) return st;
  if ((st = idwt_axis_device(ctx, v, a, vol_size, t1, 0, lo_r, hi_r, F, shift))) return st;
  return to_host(vol, dv, n * 8, ctx);
}

// Cycle spinning: the multi-level shrinkage (llr_wavelet_denoise) of every
// volume moved circularly by `shift` (sample i reads i + shift), moved back.
llr_status llr_wavelet_denoise_shifted(llr_ctx* ctx, llr_cf32* volumes, size_t n_volumes, const size_t vol_size[3],
                                       float alpha, int levels, const int shift[3]) {
  if (!ctx || !volumes || !vol_size || !shift) return fail(LLR_ERR_ARG, "null argument");
  const size_t N = prod3(vol_size);
  if (N == 0 || n_volumes == 0) return LLR_OK;
  const int L = levels < 0 ? wavelet_max_level(vol_size) : levels;
  size_t ce = 0, te = 0;
  wavelet_workspace_elems(vol_size, L, &ce, &te);
  LLR_HIP(hipSetDevice(ctx->device));
  DevBuf d, s, coef, t1, t2;
  llr_status st;
  if ((st = to_device(d, volumes, N * n_volumes * 8))) return st;
  if ((st = to_device(s, nullptr, N * n_volumes * 8))) return st;
  if ((st = to_device(coef, nullptr, ce * n_volumes * 8))) return st;
  if ((st = to_device(t1, nullptr, te * n_volumes * 8))) return st;
  if ((st = to_device(t2, nullptr, te * n_volumes * 8))) return st;
  long long fwd[3], back[3];
  for (int a = 0; a < 3; ++a) {
    back[a] = shift[a];
    fwd[a] = -(long long)shift[a];  // circshift_device: dst[i] = src[i - sh]
  }
  for (size_t v = 0; v < n_volumes; ++v)
    if ((st = circshift_device(ctx, d.as<float2>() + v * N, s.as<float2>() + v * N, 8, vol_size, fwd))) return st;
  if ((st = wavelet_denoise_device(ctx, s.p, n_volumes, vol_size, alpha, L, coef.p, t1.p, t2.p))) return st;
  for (size_t v = 0; v < n_volumes; ++v)
    if ((st = circshift_device(ctx, s.as<float2>() + v * N, d.as<float2>() + v * N, 8, vol_size, back))) return st;
  return to_host(volumes, d, N * n_volumes * 8, ctx);
}

}  // extern "C"
