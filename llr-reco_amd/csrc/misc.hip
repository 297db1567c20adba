// Element-wise and reduction kernels of the LLR path on gfx950: energy and
// sampling statistics (lib.rs:409-417), hard projection (lib.rs:392-406),
// per-volume arg-max (cufft.rs:329-337), under-sampling (data_import.rs:
// 416-438), fftshift/phase_shift3 (fftshift.rs), and the small BLAS pieces of
// the host drop-in API (cublas.rs:12-127, kernels.cu:11-67,
// dti_subspace.rs:211-221). Reductions are two-level with a fixed order, so
// repeated runs give identical sums.
#include <hip/hip_runtime.h>

#include <algorithm>

#include <cmath>

#include "common.h"
#include "internal.h"

namespace {

constexpr int RB = 256;  // reduction block

__device__ __forceinline__ double block_sum(double v, double* sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sh[i];
  __syncthreads();
  return s;
}

// partials[2*b] = sum |y|^2, partials[2*b+1] = count of nonzero, over a
// contiguous chunk per block
__global__ void __launch_bounds__(RB) k_energy(const float2* __restrict__ y, long long n, double* partials, long long chunk) {
  __shared__ double sh[RB / 64];
  long long lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
  double e = 0.0, c = 0.0;
  for (long long i = lo + threadIdx.x; i < hi; i += RB) {
    float2 v = y[i];
    e += (double)(v.x * v.x + v.y * v.y);
    c += (v.x != 0.f || v.y != 0.f) ? 1.0 : 0.0;
  }
  e = block_sum(e, sh);
  c = block_sum(c, sh);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = e;
    partials[2 * blockIdx.x + 1] = c;
  }
}

// out[j] = sum_i in[i*stride + j] over i in [0, n) for j in [0, stride):
// strided partial vectors reduced in index order by one block per j.
__global__ void __launch_bounds__(RB) k_reduce(const double* __restrict__ in, long long n, long long stride, double* out) {
  __shared__ double sh[RB / 64];
  long long j = blockIdx.x;
  double s = 0.0;
  // fixed summation order: thread t takes i = t, t + RB, ...
  for (long long i = threadIdx.x; i < n; i += RB) s += in[i * stride + j];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) out[j] = s;
}

// out[*at] = sum of in[0..n) in the fixed order of k_reduce (the index is read
// on the device, so the launch can sit in a replayed graph)
__global__ void __launch_bounds__(RB) k_reduce_at(const double* __restrict__ in, long long n, double* out,
                                                  const unsigned* at) {
  __shared__ double sh[RB / 64];
  double s = 0.0;
  for (long long i = threadIdx.x; i < n; i += RB) s += in[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) out[*at] = s;
}

// *c += 1 (the device iteration counter of a replayed graph)
__global__ void k_advance(unsigned* c) {
  if (threadIdx.x == 0) *c += 1u;
}

// out[b] = sum of in[b*chunk .. (b+1)*chunk) in a fixed order
__global__ void __launch_bounds__(RB) k_reduce_chunks(const double* __restrict__ in, long long n, long long chunk,
                                                      double* out) {
  __shared__ double sh[RB / 64];
  long long lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
  double s = 0.0;
  for (long long i = lo + threadIdx.x; i < hi; i += RB) s += in[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

__global__ void __launch_bounds__(RB) k_hard_project(float2* __restrict__ x, const float2* __restrict__ y, long long n,
                                                     double* partials, long long chunk) {
  __shared__ double sh[RB / 64];
  long long lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
  double e = 0.0;
  for (long long i = lo + threadIdx.x; i < hi; i += RB) {
    float2 m = y[i];
    if (m.x != 0.f || m.y != 0.f) {
      float2 d = csub(x[i], m);
      e += (double)(d.x * d.x + d.y * d.y);
      x[i] = m;
    }
  }
  e = block_sum(e, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = e;
}

// One pass over the measurements for the set-up: per block, the sum of |y|^2
// and the count of nonzero samples (partials[2b], partials[2b + 1], b =
// volume * gridDim.x + blockIdx.x) and the per-volume arg-max key of k_argmax.
// Four independent loads per thread and step keep enough bytes in flight.
__global__ void __launch_bounds__(RB) k_energy_argmax(const float2* __restrict__ y, long long N, double* partials,
                                                      unsigned long long* best) {
  __shared__ double sh[RB / 64];
  const long long v = blockIdx.y;
  const float2* p = y + v * N;
  const long long step = (long long)gridDim.x * RB;
  double e = 0.0, c = 0.0;
  unsigned long long b = 0;
  for (long long i0 = blockIdx.x * (long long)RB + threadIdx.x; i0 < N; i0 += 4 * step) {
    float2 s[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const long long i = i0 + u * step;
      s[u] = i < N ? p[i] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const long long i = i0 + u * step;
      const float en = s[u].x * s[u].x + s[u].y * s[u].y;
      e += (double)en;
      c += (s[u].x != 0.f || s[u].y != 0.f) ? 1.0 : 0.0;
      const unsigned long long key =
          ((unsigned long long)__float_as_uint(en) << 32) | (unsigned long long)(0xFFFFFFFFu - (unsigned)i);
      if (i < N && key > b) b = key;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long o = __shfl_down(b, off, 64);
    b = o > b ? o : b;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(best + v, b);
  e = block_sum(e, sh);
  c = block_sum(c, sh);
  if (threadIdx.x == 0) {
    const long long pb = v * gridDim.x + blockIdx.x;
    partials[2 * pb] = e;
    partials[2 * pb + 1] = c;
  }
}

// arg-max of |y|^2 per volume: pack (value bits, ~index) so that max picks the
// largest value and, among equal values, the lowest index.
__global__ void __launch_bounds__(RB) k_argmax(const float2* __restrict__ y, long long N, unsigned long long* best) {
  long long v = blockIdx.y;
  const float2* p = y + v * N;
  unsigned long long b = 0;
  for (long long i = blockIdx.x * (long long)RB + threadIdx.x; i < N; i += (long long)gridDim.x * RB) {
    float2 s = p[i];
    float e = s.x * s.x + s.y * s.y;
    unsigned long long key = ((unsigned long long)__float_as_uint(e) << 32) | (unsigned long long)(0xFFFFFFFFu - (unsigned)i);
    b = key > b ? key : b;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long o = __shfl_down(b, off, 64);
    b = o > b ? o : b;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(best + v, b);
}

__global__ void k_shrink_all(float2* x, long long n, float alpha) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    float2 v = x[i];
    float m = sqrtf(v.x * v.x + v.y * v.y);
    float s = (m == 0.f) ? 0.f : fmaxf(m - alpha, 0.f) / m;
    x[i] = make_float2(v.x * s, v.y * s);
  }
}

// Signal-model ramp exp(+i 2 pi frac((i - n/2)(c + n/2)/n)) of every volume
// along one axis, with the integer k-space correction c found by the arg-max
// (argmax_volumes packs ~index in the low 32 bits): c = s_axis - (n + 1)/2.
// (i - n/2)(c + n/2)/n = (2i - n)(2c + n) / (4n) is reduced exactly in integers
// and the angle taken with FP64 sincospi, so every entry is the FP32 rounding
// of the exact value (as the long-double host tables).
__global__ void __launch_bounds__(256) k_ramp_argmax(float2* __restrict__ tab, int n, long long stride, int nax,
                                                     const unsigned long long* __restrict__ best, int nvols) {
  const int v = blockIdx.y;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const unsigned long long idx = 0xFFFFFFFFull - (best[v] & 0xFFFFFFFFull);
    const long long sa = (long long)((idx / (unsigned long long)stride) % (unsigned long long)nax);
    const long long c = sa - ((long long)nax + 1) / 2;
    const long long M = 4LL * n;
    long long m = ((2LL * i - n) * (2 * c + n)) % M;
    if (m < 0) m += M;
    double sn, cs;
    sincospi(2.0 * (double)m / (double)M, &sn, &cs);
    tab[(long long)v * n + i] = make_float2((float)cs, (float)sn);
  }
}

// One complex value to a half pair. KEEP_NZ: a nonzero value whose two parts
// both round to zero is stored as the smallest subnormal half (2^-24, sign
// kept) in its larger part, so a sample that the hard projection treats as
// measured (y != 0, lib.rs:396) stays measured in half storage.
template <bool KEEP_NZ>
__device__ __forceinline__ unsigned pack_h16(float re, float im) {
  unsigned u = StoreF16::pack(make_float2(re, im));
  if (KEEP_NZ && (u & 0x7FFF7FFFu) == 0u && (re != 0.f || im != 0.f)) {
    const bool use_re = fabsf(re) >= fabsf(im);
    const float v = use_re ? re : im;
    const unsigned h = (v < 0.f ? 0x8001u : 0x0001u);
    u = use_re ? h : (h << 16);
  }
  return u;
}

// complex64 <-> half-precision-pair storage, scaled by s on the way (two values per thread)
template <bool KEEP_NZ>
__global__ void __launch_bounds__(256) k_to_h16(const float4* __restrict__ src, uint2* __restrict__ dst, long long n2,
                                                 float s) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n2; i += (long long)gridDim.x * blockDim.x) {
    float4 v = src[i];
    dst[i] = make_uint2(pack_h16<KEEP_NZ>(v.x * s, v.y * s), pack_h16<KEEP_NZ>(v.z * s, v.w * s));
  }
}

__global__ void __launch_bounds__(256) k_from_h16(const uint2* __restrict__ src, float4* __restrict__ dst, long long n2,
                                                   float s) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n2; i += (long long)gridDim.x * blockDim.x) {
    uint2 u = src[i];
    float2 a = StoreF16::unpack(u.x), b = StoreF16::unpack(u.y);
    dst[i] = make_float4(a.x * s, a.y * s, b.x * s, b.y * s);
  }
}

// Momentum on subspace coefficients: c = C; C = (1 + beta) c - beta P; P = c
// (first: P = C only), two complex values per thread
__global__ void __launch_bounds__(256) k_momentum(float4* __restrict__ C, float4* __restrict__ P, long long n2,
                                                   float beta, int first) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n2; i += (long long)gridDim.x * blockDim.x) {
    const float4 c = C[i];
    if (!first) {
      const float4 q = P[i];
      const float a = 1.f + beta;
      C[i] = make_float4(a * c.x - beta * q.x, a * c.y - beta * q.y, a * c.z - beta * q.z, a * c.w - beta * q.w);
    }
    P[i] = c;
  }
}

__global__ void k_under_sample(float2* data, long long nx, long long pe, long long nq, const float2* mask) {
  long long total = nx * pe * nq;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total; i += (long long)gridDim.x * blockDim.x) {
    long long line = i / nx;
    float2 m = mask[line];
    if (m.x == 0.f && m.y == 0.f) data[i] = make_float2(0.f, 0.f);
  }
}

template <class E>
__global__ void k_circshift(const E* __restrict__ src, E* __restrict__ dst, long long d0, long long d1, long long d2,
                            long long s0, long long s1, long long s2) {
  long long total = d0 * d1 * d2;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total; i += (long long)gridDim.x * blockDim.x) {
    long long x = i % d0, y = (i / d0) % d1, z = i / (d0 * d1);
    long long sx = ((x - s0) % d0 + d0) % d0, sy = ((y - s1) % d1 + d1) % d1, sz = ((z - s2) % d2 + d2) % d2;
    dst[i] = src[sx + d0 * (sy + d1 * sz)];
  }
}

__global__ void k_circshift_bytes(const unsigned char* __restrict__ src, unsigned char* __restrict__ dst, long long d0,
                                  long long d1, long long d2, long long s0, long long s1, long long s2, long long eb) {
  long long total = d0 * d1 * d2 * eb;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total; i += (long long)gridDim.x * blockDim.x) {
    long long byte = i % eb, el = i / eb;
    long long x = el % d0, y = (el / d0) % d1, z = el / (d0 * d1);
    long long sx = ((x - s0) % d0 + d0) % d0, sy = ((y - s1) % d1 + d1) % d1, sz = ((z - s2) % d2 + d2) % d2;
    dst[i] = src[(sx + d0 * (sy + d1 * sz)) * eb + byte];
  }
}

// fftshift.rs:42-74: phase pi*(i - n/2) per axis in f32, summed x + y + z
__global__ void k_phase_shift3(float2* x, long long d0, long long d1, long long d2, float sign) {
  const float PI_F = 3.14159265358979323846f;
  long long total = d0 * d1 * d2;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total; i += (long long)gridDim.x * blockDim.x) {
    long long ix = i % d0, iy = (i / d0) % d1, iz = i / (d0 * d1);
    float zs = PI_F * ((float)iz - ((float)d2 / 2.f));
    float ys = PI_F * ((float)iy - ((float)d1 / 2.f));
    float xs = PI_F * ((float)ix - ((float)d0 / 2.f));
    float tot = sign * (xs + ys + zs);
    float s, c;
    sincosf(tot, &s, &c);
    x[i] = cmul(x[i], make_float2(c, s));
  }
}

__global__ void k_scale(float2* x, long long n, float s) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    x[i] = cscale(x[i], s);
}

// kernels.cu:11-44 semantics: A[:, j, b] *= d[j, b]
__global__ void k_cdgmm(float2* a, const float* d, long long m, long long n, long long batch) {
  long long total = m * n * batch;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total; i += (long long)gridDim.x * blockDim.x) {
    long long col = (i / m) % n, b = i / (m * n);
    a[i] = cscale(a[i], d[b * n + col]);
  }
}

// cublasCgemmStridedBatched semantics (cublas.rs:12-105), one thread per C entry
__global__ void k_cgemm(long long m, long long k, long long n, long long batch, int bcast, float2 alpha, float2 beta,
                        int ta, int tb, const float2* __restrict__ A, const float2* __restrict__ B, float2* C) {
  long long lda = ta ? k : m, ldb = tb ? n : k;
  long long sa = bcast ? 0 : m * k, sb = k * n, sc = m * n;
  long long total = m * n * batch;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x) {
    long long i = e % m, j = (e / m) % n, b = e / (m * n);
    const float2* a = A + b * sa;
    const float2* bb = B + b * sb;
    float2 acc = make_float2(0.f, 0.f);
    for (long long l = 0; l < k; ++l) {
      float2 av = ta ? cconj(a[l + i * lda]) : a[i + l * lda];
      float2 bv = tb ? cconj(bb[j + l * ldb]) : bb[l + j * ldb];
      acc = cadd(acc, cmul(av, bv));
    }
    float2 r = cmul(alpha, acc);
    // like BLAS, C is not read when beta == 0 (it may hold uninitialised memory)
    if (beta.x != 0.f || beta.y != 0.f) r = cadd(r, cmul(beta, C[b * sc + i + j * m]));
    C[b * sc + i + j * m] = r;
  }
}

// out (n x m) = in (m x n)^H (or ^T) per matrix
__global__ void k_transpose(const float2* __restrict__ in, float2* __restrict__ out, long long m, long long n, long long batch,
                            int conj) {
  long long total = m * n * batch;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x) {
    long long i = e % m, j = (e / m) % n, b = e / (m * n);
    float2 v = in[e];
    if (conj) v.y = -v.y;
    out[b * m * n + i * n + j] = v;
  }
}

}  // namespace

namespace llr {

llr_status ramp_from_argmax_device(llr_ctx* ctx, float2* tab, const size_t* dims, int axis, size_t nvols,
                                   const unsigned long long* best) {
  const long long stride = axis == 0 ? 1 : axis == 1 ? (long long)dims[0] : (long long)(dims[0] * dims[1]);
  const int n = (int)dims[axis];
  hipLaunchKernelGGL(k_ramp_argmax, dim3((unsigned)((n + 255) / 256), (unsigned)nvols), dim3(256), 0, ctx->stream,
                     tab, n, stride, n, best, (int)nvols);
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

llr_status to_h16_device(llr_ctx* ctx, const float2* src, void* dst, size_t n, float s, bool keep_nonzero) {
  if (n % 2) return fail(LLR_ERR_UNSUPPORTED, "half-precision conversion needs an even number of values");
  if (keep_nonzero)
    hipLaunchKernelGGL(k_to_h16<true>, dim3(grid_for(n / 2, 256, 65536u)), dim3(256), 0, ctx->stream,
                       (const float4*)src, (uint2*)dst, (long long)(n / 2), s);
  else
    hipLaunchKernelGGL(k_to_h16<false>, dim3(grid_for(n / 2, 256, 65536u)), dim3(256), 0, ctx->stream,
                       (const float4*)src, (uint2*)dst, (long long)(n / 2), s);
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

llr_status momentum_device(llr_ctx* ctx, float2* C, float2* P, size_t n, float beta, bool first) {
  if (n % 2) return fail(LLR_ERR_UNSUPPORTED, "momentum needs an even number of coefficients");
  hipLaunchKernelGGL(k_momentum, dim3(grid_for(n / 2, 256, 65536u)), dim3(256), 0, ctx->stream, (float4*)C, (float4*)P,
                     (long long)(n / 2), beta, first ? 1 : 0);
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

llr_status from_h16_device(llr_ctx* ctx, const void* src, float2* dst, size_t n, float s) {
  if (n % 2) return fail(LLR_ERR_UNSUPPORTED, "half-precision conversion needs an even number of values");
  hipLaunchKernelGGL(k_from_h16, dim3(grid_for(n / 2, 256, 65536u)), dim3(256), 0, ctx->stream, (const uint2*)src,
                     (float4*)dst, (long long)(n / 2), s);
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

llr_status shrink_all_device(llr_ctx* ctx, float2* x, size_t n, float alpha) {
  hipLaunchKernelGGL(k_shrink_all, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, x, (long long)n, alpha);
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

// nparts blocks, each reducing a contiguous chunk; writes 2*nparts doubles
llr_status energy_partials(llr_ctx* ctx, const float2* y, size_t n, double* partials, size_t nparts) {
  long long chunk = (long long)((n + nparts - 1) / nparts);
  hipLaunchKernelGGL(k_energy, dim3((unsigned)nparts), dim3(RB), 0, ctx->stream, y, (long long)n, partials, chunk);
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

// out[0] = sum of nparts doubles (two levels through tmp when large)
llr_status reduce_sum(llr_ctx* ctx, const double* partials, size_t nparts, double* out, double* tmp, size_t tmp_n) {
  (void)tmp;
  (void)tmp_n;
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(RB), 0, ctx->stream, partials, (long long)nparts, 1LL, out);
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

llr_status reduce_strided(llr_ctx* ctx, const double* partials, size_t n, size_t stride, double* out) {
  if (stride == 1 && n > 8192) {
    // two fixed-order levels: 256 chunk sums, then their sum
    const size_t nb = 256;
    double* tmp = nullptr;
    llr_status st = red_scratch(ctx, nb, &tmp);
    if (st) return st;
    long long chunk = (long long)((n + nb - 1) / nb);
    hipLaunchKernelGGL(k_reduce_chunks, dim3((unsigned)nb), dim3(RB), 0, ctx->stream, partials, (long long)n, chunk,
                       tmp);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(RB), 0, ctx->stream, tmp, (long long)nb, 1LL, out);
    LLR_HIP(hipGetLastError());
    return LLR_OK;
  }
  hipLaunchKernelGGL(k_reduce, dim3((unsigned)stride), dim3(RB), 0, ctx->stream, partials, (long long)n,
                     (long long)stride, out);
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

// Sum of n partials into out[*at]; tmp holds 256 doubles for the two-level
// order used by reduce_strided when n > 8192 (same summation order).
llr_status reduce_at(llr_ctx* ctx, const double* partials, size_t n, double* out, const unsigned* at, double* tmp) {
  if (n > 8192) {
    const size_t nb = 256;
    long long chunk = (long long)((n + nb - 1) / nb);
    hipLaunchKernelGGL(k_reduce_chunks, dim3((unsigned)nb), dim3(RB), 0, ctx->stream, partials, (long long)n, chunk,
                       tmp);
    hipLaunchKernelGGL(k_reduce_at, dim3(1), dim3(RB), 0, ctx->stream, tmp, (long long)nb, out, at);
  } else {
    hipLaunchKernelGGL(k_reduce_at, dim3(1), dim3(RB), 0, ctx->stream, partials, (long long)n, out, at);
  }
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

llr_status advance_counter(hipStream_t s, unsigned* c) {
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, s, c);
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

llr_status hard_project_partials(llr_ctx* ctx, float2* x, const float2* y, size_t n, double* partials, size_t nparts) {
  long long chunk = (long long)((n + nparts - 1) / nparts);
  hipLaunchKernelGGL(k_hard_project, dim3((unsigned)nparts), dim3(RB), 0, ctx->stream, x, y, (long long)n, partials,
                     chunk);
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

// best[v] must be zeroed; receives (value bits << 32 | ~index) per volume
size_t energy_argmax_parts(size_t N, size_t nvols) {
  size_t gx = std::max<size_t>(1, std::min<size_t>((N + 4 * RB - 1) / (4 * RB), 2048 / std::max<size_t>(nvols, 1)));
  return gx * nvols;
}

llr_status energy_argmax_device(llr_ctx* ctx, const float2* y, size_t N, size_t nvols, double* partials,
                                unsigned long long* best) {
  if (N > 0xFFFFFFFFull) return fail(LLR_ERR_UNSUPPORTED, "volume larger than 2^32 voxels");
  const unsigned gx = (unsigned)(energy_argmax_parts(N, nvols) / nvols);
  hipLaunchKernelGGL(k_energy_argmax, dim3(gx, (unsigned)nvols), dim3(RB), 0, ctx->stream, y, (long long)N, partials,
                     best);
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

llr_status argmax_volumes(llr_ctx* ctx, const float2* y, size_t N, size_t nvols, unsigned long long* best) {
  if (N > 0xFFFFFFFFull) return fail(LLR_ERR_UNSUPPORTED, "volume larger than 2^32 voxels");
  unsigned gx = grid_for(N, RB, 1024);
  hipLaunchKernelGGL(k_argmax, dim3(gx, (unsigned)nvols), dim3(RB), 0, ctx->stream, y, (long long)N, best);
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

llr_status under_sample_device(llr_ctx* ctx, float2* data, const size_t* dims, size_t nq, const float2* mask) {
  size_t total = dims[0] * dims[1] * dims[2] * nq;
  hipLaunchKernelGGL(k_under_sample, dim3(grid_for(total, 256)), dim3(256), 0, ctx->stream, data, (long long)dims[0],
                     (long long)(dims[1] * dims[2]), (long long)nq, mask);
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

llr_status circshift_device(llr_ctx* ctx, const void* src, void* dst, size_t eb, const size_t* dims, const long long* sh) {
  long long d0 = dims[0], d1 = dims[1], d2 = dims[2];
  size_t total = dims[0] * dims[1] * dims[2];
  hipStream_t s = ctx->stream;
  unsigned g = grid_for(total, 256);
  switch (eb) {
    case 4: hipLaunchKernelGGL(k_circshift<uint32_t>, dim3(g), dim3(256), 0, s, (const uint32_t*)src, (uint32_t*)dst, d0, d1, d2, sh[0], sh[1], sh[2]); break;
    case 8: hipLaunchKernelGGL(k_circshift<uint2>, dim3(g), dim3(256), 0, s, (const uint2*)src, (uint2*)dst, d0, d1, d2, sh[0], sh[1], sh[2]); break;
    case 16: hipLaunchKernelGGL(k_circshift<uint4>, dim3(g), dim3(256), 0, s, (const uint4*)src, (uint4*)dst, d0, d1, d2, sh[0], sh[1], sh[2]); break;
    default:
      hipLaunchKernelGGL(k_circshift_bytes, dim3(grid_for(total * eb, 256)), dim3(256), 0, s, (const unsigned char*)src,
                         (unsigned char*)dst, d0, d1, d2, sh[0], sh[1], sh[2], (long long)eb);
  }
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

llr_status phase_shift3_device(llr_ctx* ctx, float2* x, const size_t* dims, int sign) {
  size_t total = dims[0] * dims[1] * dims[2];
  hipLaunchKernelGGL(k_phase_shift3, dim3(grid_for(total, 256)), dim3(256), 0, ctx->stream, x, (long long)dims[0],
                     (long long)dims[1], (long long)dims[2], (float)sign);
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

llr_status scale_device(llr_ctx* ctx, float2* x, size_t n, float s) {
  hipLaunchKernelGGL(k_scale, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, x, (long long)n, s);
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

llr_status cdgmm_device(llr_ctx* ctx, float2* a, const float* d, size_t m, size_t n, size_t batch) {
  hipLaunchKernelGGL(k_cdgmm, dim3(grid_for(m * n * batch, 256)), dim3(256), 0, ctx->stream, a, d, (long long)m,
                     (long long)n, (long long)batch);
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

llr_status cgemm_device(llr_ctx* ctx, size_t m, size_t k, size_t n, size_t batch, bool bcast_a, float2 alpha,
                        float2 beta, bool ta, bool tb, const float2* A, const float2* B, float2* C) {
  hipLaunchKernelGGL(k_cgemm, dim3(grid_for(m * n * batch, 256)), dim3(256), 0, ctx->stream, (long long)m, (long long)k,
                     (long long)n, (long long)batch, bcast_a ? 1 : 0, alpha, beta, ta ? 1 : 0, tb ? 1 : 0, A, B, C);
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

llr_status conj_transpose_device(llr_ctx* ctx, const float2* src, float2* dst, size_t m, size_t n, size_t batch,
                                 bool conj) {
  hipLaunchKernelGGL(k_transpose, dim3(grid_for(m * n * batch, 256)), dim3(256), 0, ctx->stream, src, dst, (long long)m,
                     (long long)n, (long long)batch, conj ? 1 : 0);
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

}  // namespace llr
