// Multi-level 3-D Daubechies-2 wavelet soft threshold on gfx950 (the
// wavelet_denoise step of lib.rs:430-465).
//
// Analysis along one axis (kernels.cu:155-212 formula): a[c] = sum_k lo_d[k]
// x[sym(2c+1-k)], d[c] likewise with hi_d, symmetric (half-sample) extension
// (kernels.cu:97-105), nc = (n+3)/2 coefficients per band (kernels.cu:150-152),
// approximation in [0, nc) and detail in [nc, 2nc). Synthesis (kernels.cu:
// 215-268 formula) keeps the central samples and trims to the finer level's
// length, as pywt's waverecn does. Every volume of the batch is handled by one
// launch per axis; the soft threshold (shrink_complex, lib.rs:460-465) is fused
// into the last analysis axis: it applies to every detail coefficient and, at
// the deepest level only, to the approximation.
#include <hip/hip_runtime.h>

#include <array>
#include <cmath>

#include "common.h"
#include "internal.h"

namespace {

struct Db2F {
  float dlo[4], dhi[4], rlo[4], rhi[4];
};

struct DwtArgs {
  const float2* in;
  float2* out;
  long long d[3];   // iteration box: analysis -> input dims with axis replaced by nc; synthesis -> output dims
  long long ip[3];  // input pitches
  long long ivp;    // input volume pitch
  long long op[3];
  long long ovp;
  long long n;      // analysis: input length along axis; synthesis: 2*nc coefficients
  long long nc;
  int axis;
  int nvols;
  int shrink;       // analysis: apply soft threshold
  int keep_approx;  // do not shrink the approximation sub-box (non-deepest level)
  long long h[3];   // approximation sub-box dims (for keep_approx)
  float alpha;
  Db2F f;
};

__device__ __forceinline__ long long sym(long long i, long long n) {
  // symmetric half-sample extension; |i| never exceeds n for 4-tap filters on n >= 2
  if (i < 0) i = -i - 1;
  if (i >= n) i = 2 * n - i - 1;
  if (i < 0) i = 0;  // n == 1 edge
  return i;
}

__device__ __forceinline__ float2 shrink(float2 x, float alpha) {
  float m = sqrtf(x.x * x.x + x.y * x.y);
  if (m == 0.f) return make_float2(0.f, 0.f);
  float s = fmaxf(m - alpha, 0.f) / m;
  return make_float2(x.x * s, x.y * s);
}

__global__ void __launch_bounds__(256) k_dwt_axis(DwtArgs a) {
  long long total = a.d[0] * a.d[1] * a.d[2] * a.nvols;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x) {
    long long c0 = e % a.d[0], r = e / a.d[0];
    long long c1 = r % a.d[1];
    r /= a.d[1];
    long long c2 = r % a.d[2], v = r / a.d[2];
    long long co[3] = {c0, c1, c2};
    long long c = co[a.axis];
    long long ib = v * a.ivp;
    long long ob = v * a.ovp;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax)
      if (ax != a.axis) {
        ib += co[ax] * a.ip[ax];
        ob += co[ax] * a.op[ax];
      }
    float2 sa = make_float2(0.f, 0.f), sd = make_float2(0.f, 0.f);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      long long j = sym(2 * c + 1 - k, a.n);
      float2 x = a.in[ib + j * a.ip[a.axis]];
      sa.x += a.f.dlo[k] * x.x;
      sa.y += a.f.dlo[k] * x.y;
      sd.x += a.f.dhi[k] * x.x;
      sd.y += a.f.dhi[k] * x.y;
    }
    if (a.shrink) {
      bool approx_box = true;
#pragma unroll
      for (int ax = 0; ax < 3; ++ax)
        if (ax != a.axis && co[ax] >= a.h[ax]) approx_box = false;
      if (!(a.keep_approx && approx_box)) sa = shrink(sa, a.alpha);
      sd = shrink(sd, a.alpha);
    }
    a.out[ob + c * a.op[a.axis]] = sa;
    a.out[ob + (c + a.nc) * a.op[a.axis]] = sd;
  }
}

__global__ void __launch_bounds__(256) k_idwt_axis(DwtArgs a) {
  long long total = a.d[0] * a.d[1] * a.d[2] * a.nvols;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x) {
    long long c0 = e % a.d[0], r = e / a.d[0];
    long long c1 = r % a.d[1];
    r /= a.d[1];
    long long c2 = r % a.d[2], v = r / a.d[2];
    long long co[3] = {c0, c1, c2};
    long long s = co[a.axis];
    long long ib = v * a.ivp;
    long long ob = v * a.ovp;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax)
      if (ax != a.axis) {
        ib += co[ax] * a.ip[ax];
        ob += co[ax] * a.op[ax];
      }
    float2 acc = make_float2(0.f, 0.f);
    long long cc = s + 2;  // keep-window start F-2
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      long long idx = cc - j;
      if (idx >= 0 && idx < 2 * a.nc && (idx & 1) == 0) {
        long long t = idx >> 1;
        float2 av = a.in[ib + t * a.ip[a.axis]];
        float2 dv = a.in[ib + (t + a.nc) * a.ip[a.axis]];
        acc.x += a.f.rlo[j] * av.x + a.f.rhi[j] * dv.x;
        acc.y += a.f.rlo[j] * av.y + a.f.rhi[j] * dv.y;
      }
    }
    a.out[ob + s * a.op[a.axis]] = acc;
  }
}

Db2F db2_filters() {
  long double r3 = std::sqrt(3.0L), dd = 4.0L * std::sqrt(2.0L);
  long double h[4] = {(1 + r3) / dd, (3 + r3) / dd, (3 - r3) / dd, (1 - r3) / dd};
  Db2F f;
  for (int k = 0; k < 4; ++k) {
    f.rlo[k] = (float)h[k];
    f.dlo[k] = (float)h[3 - k];
  }
  for (int k = 0; k < 4; ++k) f.rhi[k] = (k % 2 == 0) ? f.dlo[k] : -f.dlo[k];
  for (int k = 0; k < 4; ++k) f.dhi[k] = f.rhi[3 - k];
  return f;
}

}  // namespace

namespace llr {

int wavelet_max_level(const size_t* dims) {
  size_t s = dims[0];
  if (dims[1] < s) s = dims[1];
  if (dims[2] < s) s = dims[2];
  if (s < 3) return 0;
  return (int)std::floor(std::log2((double)s / 3.0));
}

// Level geometry: in_dims[l] (input box) and cdims[l] (coefficient box).
void wavelet_geometry(const size_t* dims, int L, std::vector<std::array<size_t, 3>>& in_dims,
                      std::vector<std::array<size_t, 3>>& cdims) {
  in_dims.resize(L);
  cdims.resize(L);
  size_t cur[3] = {dims[0], dims[1], dims[2]};
  for (int l = 0; l < L; ++l) {
    for (int a = 0; a < 3; ++a) {
      in_dims[l][a] = cur[a];
      cdims[l][a] = 2 * ((cur[a] + 3) / 2);
    }
    for (int a = 0; a < 3; ++a) cur[a] = cdims[l][a] / 2;
  }
}

// Workspace elements needed per volume: coefficient boxes of every level plus
// two axis temporaries sized for level 0.
size_t wavelet_workspace_elems(const size_t* dims, int L, size_t* coef_elems, size_t* tmp_elems) {
  std::vector<std::array<size_t, 3>> in_dims, cdims;
  wavelet_geometry(dims, L, in_dims, cdims);
  size_t c = 0;
  for (int l = 0; l < L; ++l) c += cdims[l][0] * cdims[l][1] * cdims[l][2];
  size_t t = L > 0 ? cdims[0][0] * cdims[0][1] * cdims[0][2] : 0;
  *coef_elems = c;
  *tmp_elems = t;
  return c + 2 * t;
}

static hipError_t launch_dwt(DwtArgs& a, bool synth, hipStream_t s) {
  long long total = a.d[0] * a.d[1] * a.d[2] * a.nvols;
  unsigned grid = grid_for((size_t)total, 256, 1u << 20);
  if (synth)
    hipLaunchKernelGGL(k_idwt_axis, dim3(grid), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(k_dwt_axis, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

// In-place wavelet soft threshold of nvols volumes. coef: workspace of
// nvols * coef_elems, t1/t2: nvols * tmp_elems each.
llr_status wavelet_denoise_device(llr_ctx* ctx, float2* vols, size_t nvols, const size_t* dims, float alpha, int L,
                                  float2* coef, float2* t1, float2* t2) {
  Db2F f = db2_filters();
  size_t N = dims[0] * dims[1] * dims[2];
  if (L <= 0) return shrink_all_device(ctx, vols, N * nvols, alpha);  // the volume is its own approximation
  std::vector<std::array<size_t, 3>> in_dims, cdims;
  wavelet_geometry(dims, L, in_dims, cdims);
  std::vector<float2*> cb(L);
  std::vector<size_t> cvol(L);
  {
    float2* p = coef;
    for (int l = 0; l < L; ++l) {
      cvol[l] = cdims[l][0] * cdims[l][1] * cdims[l][2];
      cb[l] = p;
      p += cvol[l] * nvols;
    }
  }
  size_t tvol = cvol[0];
  hipStream_t s = ctx->stream;
  // ---------------- analysis
  for (int l = 0; l < L; ++l) {
    const float2* src;
    long long sp[3], svp;
    if (l == 0) {
      src = vols;
      sp[0] = 1;
      sp[1] = (long long)dims[0];
      sp[2] = (long long)(dims[0] * dims[1]);
      svp = (long long)N;
    } else {
      src = cb[l - 1];
      sp[0] = 1;
      sp[1] = (long long)cdims[l - 1][0];
      sp[2] = (long long)(cdims[l - 1][0] * cdims[l - 1][1]);
      svp = (long long)cvol[l - 1];
    }
    const size_t* in = in_dims[l].data();
    const size_t* cd = cdims[l].data();
    long long d1[3] = {(long long)cd[0], (long long)in[1], (long long)in[2]};
    long long d2[3] = {(long long)cd[0], (long long)cd[1], (long long)in[2]};
    long long p1[3] = {1, d1[0], d1[0] * d1[1]}, p2[3] = {1, d2[0], d2[0] * d2[1]};
    long long pc[3] = {1, (long long)cd[0], (long long)(cd[0] * cd[1])};
    for (int ax = 0; ax < 3; ++ax) {
      DwtArgs a{};
      a.f = f;
      a.nvols = (int)nvols;
      a.axis = ax;
      a.n = (long long)in[ax];
      a.nc = (long long)(cd[ax] / 2);
      if (ax == 0) {
        a.in = src;
        for (int k = 0; k < 3; ++k) a.ip[k] = sp[k];
        a.ivp = svp;
        a.out = t1;
        for (int k = 0; k < 3; ++k) a.op[k] = p1[k];
        a.ovp = (long long)tvol;
        a.d[0] = a.nc;
        a.d[1] = (long long)in[1];
        a.d[2] = (long long)in[2];
      } else if (ax == 1) {
        a.in = t1;
        for (int k = 0; k < 3; ++k) a.ip[k] = p1[k];
        a.ivp = (long long)tvol;
        a.out = t2;
        for (int k = 0; k < 3; ++k) a.op[k] = p2[k];
        a.ovp = (long long)tvol;
        a.d[0] = d1[0];
        a.d[1] = a.nc;
        a.d[2] = (long long)in[2];
      } else {
        a.in = t2;
        for (int k = 0; k < 3; ++k) a.ip[k] = p2[k];
        a.ivp = (long long)tvol;
        a.out = cb[l];
        for (int k = 0; k < 3; ++k) a.op[k] = pc[k];
        a.ovp = (long long)cvol[l];
        a.d[0] = d2[0];
        a.d[1] = d2[1];
        a.d[2] = a.nc;
        a.shrink = 1;
        a.keep_approx = (l != L - 1);
        a.h[0] = (long long)cd[0] / 2;
        a.h[1] = (long long)cd[1] / 2;
        a.h[2] = (long long)cd[2] / 2;
        a.alpha = alpha;
      }
      hipError_t e = launch_dwt(a, false, s);
      if (e != hipSuccess) return hip_fail(e, "dwt analysis launch");
    }
  }
  // ---------------- synthesis
  for (int l = L - 1; l >= 0; --l) {
    const size_t* cd = cdims[l].data();
    const size_t* od = in_dims[l].data();
    long long pc[3] = {1, (long long)cd[0], (long long)(cd[0] * cd[1])};
    long long d2[3] = {(long long)cd[0], (long long)cd[1], (long long)od[2]};
    long long d1[3] = {(long long)cd[0], (long long)od[1], (long long)od[2]};
    long long p2[3] = {1, d2[0], d2[0] * d2[1]}, p1[3] = {1, d1[0], d1[0] * d1[1]};
    for (int ax = 2; ax >= 0; --ax) {
      DwtArgs a{};
      a.f = f;
      a.nvols = (int)nvols;
      a.axis = ax;
      a.nc = (long long)(cd[ax] / 2);
      a.n = 2 * a.nc;
      if (ax == 2) {
        a.in = cb[l];
        for (int k = 0; k < 3; ++k) a.ip[k] = pc[k];
        a.ivp = (long long)cvol[l];
        a.out = t2;
        for (int k = 0; k < 3; ++k) a.op[k] = p2[k];
        a.ovp = (long long)tvol;
        for (int k = 0; k < 3; ++k) a.d[k] = d2[k];
      } else if (ax == 1) {
        a.in = t2;
        for (int k = 0; k < 3; ++k) a.ip[k] = p2[k];
        a.ivp = (long long)tvol;
        a.out = t1;
        for (int k = 0; k < 3; ++k) a.op[k] = p1[k];
        a.ovp = (long long)tvol;
        for (int k = 0; k < 3; ++k) a.d[k] = d1[k];
      } else {
        a.in = t1;
        for (int k = 0; k < 3; ++k) a.ip[k] = p1[k];
        a.ivp = (long long)tvol;
        if (l == 0) {
          a.out = vols;
          a.op[0] = 1;
          a.op[1] = (long long)dims[0];
          a.op[2] = (long long)(dims[0] * dims[1]);
          a.ovp = (long long)N;
        } else {
          a.out = cb[l - 1];
          a.op[0] = 1;
          a.op[1] = (long long)cdims[l - 1][0];
          a.op[2] = (long long)(cdims[l - 1][0] * cdims[l - 1][1]);
          a.ovp = (long long)cvol[l - 1];
        }
        a.d[0] = (long long)od[0];
        a.d[1] = (long long)od[1];
        a.d[2] = (long long)od[2];
      }
      hipError_t e = launch_dwt(a, true, s);
      if (e != hipSuccess) return hip_fail(e, "dwt synthesis launch");
    }
  }
  return LLR_OK;
}

}  // namespace llr
