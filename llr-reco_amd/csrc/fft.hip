// 3-D signal model on gfx950: LDS-staged four-step line FFTs.
//
// A line of length N = R1*R2 is transformed by one thread group of P =
// max(R1, R2) lanes: stage 1 loads x[j1*R2 + j2] straight from HBM into
// registers (one lane per j2), does an R1-point DFT over j1, multiplies the
// twiddle W_N^{j2*k1} and parks the result in LDS; stage 2 reads it back (one
// lane per k1), does an R2-point DFT over j2 and stores X[k1 + R1*k2].
// A workgroup holds T lines. For the strided passes (y, z) the T lines are
// neighbours along x, so every load/store instruction touches T contiguous
// complex values; for the x pass the lanes of one line are contiguous.
//
// The signal model's separable phase ramps (cufft.rs:290-326) are folded into
// the passes: each axis pass multiplies its input by that axis' ramp table and
// its output by the other ramp table, and applies 1/sqrt(n_axis) of the
// unitary scale (cufft.rs:161-163). The z pass of the forward model is fused
// with the hard projection (lib.rs:392-406) and the z pass of the adjoint, so a
// z line goes k-space and back without leaving the chip.
#include <hip/hip_runtime.h>

#include <cmath>

#include "common.h"
#include "internal.h"

namespace {

// exp(-2*pi*i*k/16) components, k in [0,16)
__device__ __forceinline__ constexpr float cos16(int k) {
  switch (k & 15) {
    case 0: return 1.0f;
    case 1: return 0.923879532511286756f;
    case 2: return 0.707106781186547524f;
    case 3: return 0.382683432365089772f;
    case 4: return 0.0f;
    case 5: return -0.382683432365089772f;
    case 6: return -0.707106781186547524f;
    case 7: return -0.923879532511286756f;
    case 8: return -1.0f;
    case 9: return -0.923879532511286756f;
    case 10: return -0.707106781186547524f;
    case 11: return -0.382683432365089772f;
    case 12: return 0.0f;
    case 13: return 0.382683432365089772f;
    case 14: return 0.707106781186547524f;
    default: return 0.923879532511286756f;
  }
}
__device__ __forceinline__ constexpr float sin16(int k) { return cos16(k - 4); }

// forward twiddle exp(-2*pi*i*k/R) for R | 32
template <int R>
__device__ __forceinline__ float2 tw_const(int k) {
  if constexpr (R <= 16) {
    int k16 = k * (16 / R);
    return make_float2(cos16(k16), -sin16(k16));
  } else {
    // R == 32: half steps; cos(2*pi*k/32)
    constexpr float c32[8] = {1.0f, 0.980785280403230449f, 0.923879532511286756f, 0.831469612302545237f,
                              0.707106781186547524f, 0.555570233019602225f, 0.382683432365089772f,
                              0.195090322016128268f};
    int kk = k & 31;
    // cos(2pi kk/32) with symmetry
    float c, s;
    int q = kk >> 3, r = kk & 7;
    float cr = c32[r], sr = (r == 0) ? 0.0f : c32[8 - r];
    if (q == 0) { c = cr; s = sr; }
    else if (q == 1) { c = -sr; s = cr; }
    else if (q == 2) { c = -cr; s = -sr; }
    else { c = sr; s = -cr; }
    return make_float2(c, -s);
  }
}

// In-register forward DFT of length R (natural order in and out), radix-2 DIT.
template <int R>
struct SmallDFT {
  __device__ __forceinline__ static void run(float2* v) {
    float2 e[R / 2], o[R / 2];
#pragma unroll
    for (int i = 0; i < R / 2; ++i) {
      e[i] = v[2 * i];
      o[i] = v[2 * i + 1];
    }
    SmallDFT<R / 2>::run(e);
    SmallDFT<R / 2>::run(o);
#pragma unroll
    for (int k = 0; k < R / 2; ++k) {
      float2 t = cmul(o[k], tw_const<R>(k));
      v[k] = cadd(e[k], t);
      v[k + R / 2] = csub(e[k], t);
    }
  }
};
template <>
struct SmallDFT<1> {
  __device__ __forceinline__ static void run(float2*) {}
};
template <>
struct SmallDFT<2> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};
template <>
struct SmallDFT<4> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
    float2 c = cadd(v[1], v[3]), d = csub(v[1], v[3]);
    // d * (-i)
    float2 dmi = make_float2(d.y, -d.x);
    v[0] = cadd(a, c);
    v[2] = csub(a, c);
    v[1] = cadd(b, dmi);
    v[3] = csub(b, dmi);
  }
};

// DFT with sign: forward exp(-), inverse exp(+) via conjugation symmetry.
template <int R>
__device__ __forceinline__ void dft(float2* v, bool inv) {
  if (inv) {
#pragma unroll
    for (int i = 0; i < R; ++i) v[i].y = -v[i].y;
  }
  SmallDFT<R>::run(v);
  if (inv) {
#pragma unroll
    for (int i = 0; i < R; ++i) v[i].y = -v[i].y;
  }
}

struct PassArgs {
  const float2* src;
  float2* dst;
  long long stride;   // element stride along the transformed axis
  int U;              // number of lines along the tiled (neighbour) dimension
  long long ustride;  // element stride between neighbouring lines
  long long wstride;  // element stride of the grid's y dimension
  long long vstride;  // element stride between volumes
  const float2* pre;  // [v][N] input ramp table or null
  const float2* post; // [v][N] output ramp table or null
  int pre_conj, post_conj;
  float scale;
  const float2* tw;   // exp(-2*pi*i*m/N)
  int inverse;
};

template <int R1, int R2>
struct Layout {
  static constexpr int N = R1 * R2;
  static constexpr int P = R1 > R2 ? R1 : R2;
  static constexpr int T = (256 / P) > 0 ? (256 / P) : 1;
  static constexpr int R2P = R2 + 1;
  static constexpr int LS = R1 * R2P + 1;
  static constexpr int THREADS = T * P;
};

__device__ __forceinline__ float2 ld_tab(const float2* tab, long long v, int n, int i, int conj) {
  float2 p = tab[v * n + i];
  if (conj) p.y = -p.y;
  return p;
}

// One axis pass over all lines of all volumes.
template <int R1, int R2, bool TFAST>
__global__ void __launch_bounds__((Layout<R1, R2>::THREADS)) k_fft_pass(PassArgs a) {
  using L = Layout<R1, R2>;
  constexpr int N = L::N, P = L::P, T = L::T;
  __shared__ float2 lds[T * L::LS];
  const int tid = threadIdx.x;
  const int t = TFAST ? (tid % T) : (tid / P);
  const int j = TFAST ? (tid / T) : (tid % P);  // lane within the line
  const int u = blockIdx.x * T + t;
  const long long v = blockIdx.z;
  const bool active = u < a.U;
  const long long base = v * a.vstride + (long long)blockIdx.y * a.wstride + (long long)u * a.ustride;
  const bool inv = a.inverse != 0;

  // stage 1: lane j2 = j loads x[j1*R2 + j2], R1-point DFT over j1
  if (active && j < R2) {
    float2 x[R1];
#pragma unroll
    for (int j1 = 0; j1 < R1; ++j1) {
      int idx = j1 * R2 + j;
      float2 val = a.src[base + (long long)idx * a.stride];
      if (a.pre) val = cmul(val, ld_tab(a.pre, v, N, idx, a.pre_conj));
      x[j1] = val;
    }
    dft<R1>(x, inv);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      float2 w = a.tw[j * k1];
      if (inv) w.y = -w.y;
      lds[t * L::LS + k1 * L::R2P + j] = cmul(x[k1], w);
    }
  }
  __syncthreads();
  // stage 2: lane k1 = j reads the R2 values, R2-point DFT, stores X[k1 + R1*k2]
  if (active && j < R1) {
    float2 y[R2];
#pragma unroll
    for (int j2 = 0; j2 < R2; ++j2) y[j2] = lds[t * L::LS + j * L::R2P + j2];
    dft<R2>(y, inv);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) {
      int k = j + R1 * k2;
      float2 val = cscale(y[k2], a.scale);
      if (a.post) val = cmul(val, ld_tab(a.post, v, N, k, a.post_conj));
      a.dst[base + (long long)k * a.stride] = val;
    }
  }
}

struct ZProjArgs {
  float2* data;        // image-space iterate after the forward x and y passes
  const float2* meas;  // measurements (k-space), same layout
  long long stride, ustride, wstride, vstride;
  int U;
  const float2* pk;    // [v][nz] exp(+i phi_ksp,z)
  const float2* pi;    // [v][nz] exp(+i phi_img,z)
  float scale;         // 1/sqrt(nz)
  const float2* tw;
  double* partial;     // one partial sum of |x - y|^2 per workgroup
};

// Forward z pass -> hard projection -> adjoint z pass on one tile of z lines.
// The forward output ramp conj(pi) and the adjoint input ramp pi cancel on
// unsampled entries, so the projection compares x*scale with y*pi and, where
// sampled, substitutes y*pi (|pi| = 1, so |x*conj(pi) - y| = |x - y*pi|).
template <int R1, int R2>
__global__ void __launch_bounds__((Layout<R1, R2>::THREADS)) k_fft_z_project(ZProjArgs a) {
  using L = Layout<R1, R2>;
  constexpr int N = L::N, T = L::T;
  __shared__ float2 lds[T * L::LS];
  __shared__ double red[L::THREADS / 64];
  const int tid = threadIdx.x;
  const int t = tid % T;
  const int j = tid / T;
  const int u = blockIdx.x * T + t;
  const long long v = blockIdx.z;
  const bool active = u < a.U;
  const long long base = v * a.vstride + (long long)blockIdx.y * a.wstride + (long long)u * a.ustride;
  double err = 0.0;

  if (active && j < R2) {
    float2 x[R1];
#pragma unroll
    for (int j1 = 0; j1 < R1; ++j1) {
      int idx = j1 * R2 + j;
      x[j1] = cmul(a.data[base + (long long)idx * a.stride], a.pk[v * N + idx]);
    }
    dft<R1>(x, false);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) lds[t * L::LS + k1 * L::R2P + j] = cmul(x[k1], a.tw[j * k1]);
  }
  __syncthreads();
  float2 y[R2];
  if (active && j < R1) {
#pragma unroll
    for (int j2 = 0; j2 < R2; ++j2) y[j2] = lds[t * L::LS + j * L::R2P + j2];
    dft<R2>(y, false);
    // k-space sample k = j + R1*k2: project, then start the adjoint (R2-point
    // inverse DFT over k2 for the fixed k1 = j)
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) {
      int k = j + R1 * k2;
      float2 xv = cscale(y[k2], a.scale);
      float2 m = a.meas[base + (long long)k * a.stride];
      if (m.x != 0.0f || m.y != 0.0f) {
        float2 mp = cmul(m, a.pi[v * N + k]);
        float2 d = csub(xv, mp);
        err += (double)cnorm(d);
        xv = mp;
      }
      y[k2] = xv;
    }
    dft<R2>(y, true);
  }
  __syncthreads();  // every stage-2 read of lds is done
  if (active && j < R1) {
#pragma unroll
    for (int j2 = 0; j2 < R2; ++j2) {
      float2 w = a.tw[j * j2];
      w.y = -w.y;
      lds[t * L::LS + j * L::R2P + j2] = cmul(y[j2], w);
    }
  }
  __syncthreads();
  if (active && j < R2) {
    float2 x[R1];
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) x[k1] = lds[t * L::LS + k1 * L::R2P + j];
    dft<R1>(x, true);
#pragma unroll
    for (int j1 = 0; j1 < R1; ++j1) {
      int idx = j + R2 * j1;
      float2 pk = a.pk[v * N + idx];
      a.data[base + (long long)idx * a.stride] = cscale(cmulc(x[j1], pk), a.scale);
    }
  }
  // workgroup reduction of err (fixed order: deterministic)
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) err += __shfl_down(err, off, 64);
  if ((tid & 63) == 0) red[tid >> 6] = err;
  __syncthreads();
  if (tid == 0) {
    double s = 0.0;
    for (int w = 0; w < L::THREADS / 64; ++w) s += red[w];
    size_t bid = ((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    a.partial[bid] = s;
  }
}

template <int R1, int R2>
hipError_t launch_pass(const PassArgs& a, long long W, long long nvols, bool tfast, hipStream_t s) {
  using L = Layout<R1, R2>;
  dim3 grid((a.U + L::T - 1) / L::T, (unsigned)W, (unsigned)nvols);
  if (tfast)
    hipLaunchKernelGGL((k_fft_pass<R1, R2, true>), grid, dim3(L::THREADS), 0, s, a);
  else
    hipLaunchKernelGGL((k_fft_pass<R1, R2, false>), grid, dim3(L::THREADS), 0, s, a);
  return hipGetLastError();
}

template <int R1, int R2>
hipError_t launch_zproj(const ZProjArgs& a, long long W, long long nvols, hipStream_t s, size_t* nparts) {
  using L = Layout<R1, R2>;
  dim3 grid((a.U + L::T - 1) / L::T, (unsigned)W, (unsigned)nvols);
  *nparts = (size_t)grid.x * grid.y * grid.z;
  if (a.partial == nullptr) return hipSuccess;  // size query
  hipLaunchKernelGGL((k_fft_z_project<R1, R2>), grid, dim3(L::THREADS), 0, s, a);
  return hipGetLastError();
}

}  // namespace

namespace llr {

bool fft_size_supported(size_t n) {
  return n == 1 || n == 2 || n == 4 || n == 8 || n == 16 || n == 32 || n == 64 || n == 128 || n == 256 ||
         n == 512;
}

// Dispatch an axis pass for line length n.
static hipError_t dispatch_pass(int n, const PassArgs& a, long long W, long long nv, bool tfast, hipStream_t s) {
  switch (n) {
    case 2: return launch_pass<2, 1>(a, W, nv, tfast, s);
    case 4: return launch_pass<2, 2>(a, W, nv, tfast, s);
    case 8: return launch_pass<4, 2>(a, W, nv, tfast, s);
    case 16: return launch_pass<4, 4>(a, W, nv, tfast, s);
    case 32: return launch_pass<8, 4>(a, W, nv, tfast, s);
    case 64: return launch_pass<8, 8>(a, W, nv, tfast, s);
    case 128: return launch_pass<16, 8>(a, W, nv, tfast, s);
    case 256: return launch_pass<16, 16>(a, W, nv, tfast, s);
    case 512: return launch_pass<32, 16>(a, W, nv, tfast, s);
    default: return hipErrorInvalidValue;
  }
}

// One pass along `axis` of nvols volumes of size dims; src may equal dst.
llr_status fft_axis_pass(llr_ctx* ctx, const float2* src, float2* dst, const size_t* dims, size_t nvols,
                         int axis, bool inverse, const float2* pre, bool pre_conj, const float2* post,
                         bool post_conj, float scale) {
  int n = (int)dims[axis];
  if (n == 1 && !pre && !post && scale == 1.0f && src == dst) return LLR_OK;
  if (!fft_size_supported((size_t)n) && n != 1)
    return fail(LLR_ERR_UNSUPPORTED, "FFT length " + std::to_string(n) + " is not a supported power of two (2..512)");
  const float2* tw = nullptr;
  if (n > 1) {
    llr_status st = get_twiddles(ctx, n, &tw);
    if (st) return st;
  }
  long long nx = dims[0], ny = dims[1], nz = dims[2];
  PassArgs a{};
  a.src = src;
  a.dst = dst;
  a.vstride = nx * ny * nz;
  a.pre = pre;
  a.post = post;
  a.pre_conj = pre_conj;
  a.post_conj = post_conj;
  a.scale = scale;
  a.tw = tw;
  a.inverse = inverse;
  long long W;
  bool tfast;
  if (axis == 0) {
    a.stride = 1;
    a.U = (int)ny;
    a.ustride = nx;
    W = nz;
    a.wstride = nx * ny;
    tfast = false;
  } else if (axis == 1) {
    a.stride = nx;
    a.U = (int)nx;
    a.ustride = 1;
    W = nz;
    a.wstride = nx * ny;
    tfast = true;
  } else {
    a.stride = nx * ny;
    a.U = (int)nx;
    a.ustride = 1;
    W = ny;
    a.wstride = nx;
    tfast = true;
  }
  if (n == 1) {
    // a length-1 "transform" is the ramp/scale only: run it as a 2-point-free
    // elementwise pass through the x-pass kernel of length 1 is not provided;
    // handled by the caller (dims of 1 only occur in tests with pre/post null).
    return fail(LLR_ERR_UNSUPPORTED, "length-1 axis with phase ramp");
  }
  hipError_t e = dispatch_pass(n, a, W, (long long)nvols, tfast, ctx->stream);
  if (e != hipSuccess) return hip_fail(e, "fft axis pass launch");
  return LLR_OK;
}

// Fused forward-z / hard projection / adjoint-z. Writes one partial error sum
// per workgroup into `partial` (size from fft_zproj_parts).
llr_status fft_zproj(llr_ctx* ctx, float2* data, const float2* meas, const size_t* dims, size_t nvols,
                     const float2* pk, const float2* pi, double* partial, size_t* nparts) {
  int n = (int)dims[2];
  const float2* tw = nullptr;
  if (partial) {
    llr_status st = get_twiddles(ctx, n, &tw);
    if (st) return st;
  }
  long long nx = dims[0], ny = dims[1], nz = dims[2];
  ZProjArgs a{};
  a.data = data;
  a.meas = meas;
  a.stride = nx * ny;
  a.U = (int)nx;
  a.ustride = 1;
  a.wstride = nx;
  a.vstride = nx * ny * nz;
  a.pk = pk;
  a.pi = pi;
  a.scale = (float)(1.0 / std::sqrt((double)nz));
  a.tw = tw;
  a.partial = partial;
  hipError_t e;
  switch (n) {
    case 2: e = launch_zproj<2, 1>(a, ny, nvols, ctx->stream, nparts); break;
    case 4: e = launch_zproj<2, 2>(a, ny, nvols, ctx->stream, nparts); break;
    case 8: e = launch_zproj<4, 2>(a, ny, nvols, ctx->stream, nparts); break;
    case 16: e = launch_zproj<4, 4>(a, ny, nvols, ctx->stream, nparts); break;
    case 32: e = launch_zproj<8, 4>(a, ny, nvols, ctx->stream, nparts); break;
    case 64: e = launch_zproj<8, 8>(a, ny, nvols, ctx->stream, nparts); break;
    case 128: e = launch_zproj<16, 8>(a, ny, nvols, ctx->stream, nparts); break;
    case 256: e = launch_zproj<16, 16>(a, ny, nvols, ctx->stream, nparts); break;
    case 512: e = launch_zproj<32, 16>(a, ny, nvols, ctx->stream, nparts); break;
    default: return fail(LLR_ERR_UNSUPPORTED, "z length " + std::to_string(n) + " is not a supported power of two");
  }
  if (e != hipSuccess) return hip_fail(e, "fft z projection launch");
  return LLR_OK;
}

}  // namespace llr
