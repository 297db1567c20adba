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

// exp(-2*pi*i*k/12) components, k in [0,12)
__device__ __forceinline__ constexpr float cos12(int k) {
  switch (((k % 12) + 12) % 12) {
    case 0: return 1.0f;
    case 1: return 0.866025403784438647f;
    case 2: return 0.5f;
    case 3: return 0.0f;
    case 4: return -0.5f;
    case 5: return -0.866025403784438647f;
    case 6: return -1.0f;
    case 7: return -0.866025403784438647f;
    case 8: return -0.5f;
    case 9: return 0.0f;
    case 10: return 0.5f;
    default: return 0.866025403784438647f;
  }
}
__device__ __forceinline__ constexpr float sin12(int k) { return cos12(k - 3); }

// forward twiddle exp(-2*pi*i*k/R) for R | 32 or R | 12
template <int R>
__device__ __forceinline__ float2 tw_const(int k) {
  if constexpr (12 % R == 0 && 16 % R != 0) {
    int k12 = k * (12 / R);
    return make_float2(cos12(k12), -sin12(k12));
  } else if constexpr (R <= 16) {
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
      // W^0 = 1 and W^(R/4) = -i taken exactly (k is a constant once unrolled:
      // x * 0 does not fold without finite-math, so cmul would spend a multiply
      // and an add on each)
      const float2 t = k == 0 ? o[k] : (4 * k == R ? make_float2(o[k].y, -o[k].x) : cmul(o[k], tw_const<R>(k)));
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
struct SmallDFT<3> {
  __device__ __forceinline__ static void run(float2* v) {
    const float h = 0.866025403784438647f;  // sqrt(3)/2
    float2 a = v[0], t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
    float2 m = make_float2(a.x - 0.5f * t1.x, a.y - 0.5f * t1.y);
    float2 r = make_float2(h * t2.y, -h * t2.x);  // -i * (sqrt(3)/2) * t2
    v[0] = cadd(a, t1);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
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
  const void* src;    // complex values in the pass's storage (ST::T)
  void* dst;
  int stride;         // element stride along the transformed axis
  int U;              // number of lines along the tiled (neighbour) dimension
  int ustride;        // element stride between neighbouring lines
  long long wstride;  // element stride of the grid's y dimension
  long long vstride;  // element stride between volumes
  const float2* pre;  // [v][N] input ramp table or null
  const float2* post; // [v][N] output ramp table or null
  int pre_conj, post_conj;
  float scale;
  const float2* tw;   // exp(-2*pi*i*m/N)
  int tiles;          // tiles of T lines along U
  int tiles_per_wg;
  unsigned* flag;     // line check (x passes): set to 1 when a line is partly measured
};

// The value a stored complex holds is nonzero (the hard projection's
// measured-sample test, applied to what a later pass will read back).
template <class SO>
__device__ __forceinline__ bool stored_nonzero(float2 v) {
  if constexpr (SO::bytes == 4) return (SO::pack(v) & 0x7fff7fffu) != 0u;
  return v.x != 0.f || v.y != 0.f;
}

constexpr int pow2_floor(int x) { return x >= 256 ? 256 : x >= 128 ? 128 : x >= 64 ? 64 : x >= 32 ? 32 : x >= 16 ? 16 : x >= 8 ? 8 : x >= 4 ? 4 : x >= 2 ? 2 : 1; }

// MINT: at least this many lines per workgroup while it keeps 512 threads or
// fewer (half storage takes 32 so a strided pass moves 128-byte runs).
template <int R1, int R2, int MINT = 1>
struct Layout {
  static constexpr int N = R1 * R2;
  static constexpr int P = R1 > R2 ? R1 : R2;
  static constexpr int T0 = pow2_floor(256 / P);
  static constexpr int TM = MINT < 512 / P ? MINT : 512 / P;
  static constexpr int T = T0 > TM ? T0 : TM;  // lines per workgroup (power of two: aligned x runs)
  static constexpr int R2P = R2 + 1;
  static constexpr int LS = R1 * R2P + 1;
  static constexpr int THREADS = T * P;
};

// One axis pass. Each workgroup owns `tiles_per_wg` consecutive tiles of T
// lines of one (w, volume) row: the twiddles and ramp tables (with the
// direction, conjugation and 1/sqrt(n) folded in) are staged in LDS once, and
// the next tile's loads are issued before the current tile's second stage so
// they overlap its compute and stores. Offsets inside a volume are 32-bit.
// SO: storage of the output (the input's by default). CHECK (lines along the
// contiguous axis only): count each line's nonzero inputs and stored outputs
// and set *a.flag when either count is neither 0 nor N.
template <int R1, int R2, bool TFAST, bool INV, bool RAMPS, class ST, int MINT, class SO = ST, bool CHECK = false>
__global__ void __launch_bounds__((Layout<R1, R2, MINT>::THREADS)) k_fft_pass(PassArgs a) {
  using L = Layout<R1, R2, MINT>;
  using E = typename ST::T;
  using EO = typename SO::T;
  constexpr int N = L::N, P = L::P, T = L::T, TH = L::THREADS;
  static_assert(!CHECK || !TFAST, "line checks are for contiguous lines");
  __shared__ float2 lds[T * L::LS];
  __shared__ float2 s_tw[N];
  __shared__ float2 s_pre[RAMPS ? N : 1];
  __shared__ float2 s_post[RAMPS ? N : 1];
  __shared__ int s_cnt[CHECK ? 2 * 2 * T : 1];  // [tile parity][in, out][line]
  const int tid = threadIdx.x;
  const int t = TFAST ? (tid % T) : (tid / P);
  const int j = TFAST ? (tid / T) : (tid % P);
  const long long v = blockIdx.z;
  for (int i = tid; i < N; i += TH) {
    float2 w = a.tw[i];
    s_tw[i] = INV ? cconj(w) : w;
    if constexpr (RAMPS) {
      float2 p = a.pre[v * N + i];
      s_pre[i] = a.pre_conj ? cconj(p) : p;
      float2 q = a.post[v * N + i];
      s_post[i] = cscale(a.post_conj ? cconj(q) : q, a.scale);
    }
  }
  if constexpr (CHECK)
    for (int i = tid; i < 4 * T; i += TH) s_cnt[i] = 0;
  const long long row = v * a.vstride + (long long)blockIdx.y * a.wstride;
  const E* __restrict__ src = (const E*)a.src + row;
  EO* __restrict__ dst = (EO*)a.dst + row;
  const int tile0 = blockIdx.x * a.tiles_per_wg;
  const int ntiles = min(a.tiles_per_wg, a.tiles - tile0);
  float2 xn[R1];
  auto load = [&](int tile, float2* x) {
    int u = tile * T + t;
    if (u < a.U && j < R2) {
      const E* p = src + (unsigned)(u * a.ustride + j * a.stride);
      unsigned step = (unsigned)(R2 * a.stride);
#pragma unroll
      for (int j1 = 0; j1 < R1; ++j1) x[j1] = ST::ldnt(p + j1 * step);
    }
  };
  load(tile0, xn);
  __syncthreads();  // tables ready
  for (int it = 0; it < ntiles; ++it) {
    const int u = (tile0 + it) * T + t;
    const bool active = u < a.U;
    int* cnt = s_cnt + (CHECK ? (it & 1) * 2 * T : 0);
    if (active && j < R2) {
      if constexpr (CHECK) {
        int c = 0;
#pragma unroll
        for (int j1 = 0; j1 < R1; ++j1) c += (xn[j1].x != 0.f || xn[j1].y != 0.f) ? 1 : 0;
        atomicAdd(&cnt[t], c);
      }
      float2 x[R1];
#pragma unroll
      for (int j1 = 0; j1 < R1; ++j1) x[j1] = RAMPS ? cmul(xn[j1], s_pre[j1 * R2 + j]) : xn[j1];
      if (INV) {
#pragma unroll
        for (int q = 0; q < R1; ++q) x[q].y = -x[q].y;
      }
      SmallDFT<R1>::run(x);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) {
        float2 y = x[k1];
        if (INV) y.y = -y.y;
        lds[t * L::LS + k1 * L::R2P + j] = cmul(y, s_tw[j * k1]);
      }
    }
    __syncthreads();
    if (it + 1 < ntiles) load(tile0 + it + 1, xn);
    if (active && j < R1) {
      float2 y[R2];
#pragma unroll
      for (int j2 = 0; j2 < R2; ++j2) {
        y[j2] = lds[t * L::LS + j * L::R2P + j2];
        if (INV) y[j2].y = -y[j2].y;
      }
      SmallDFT<R2>::run(y);
      EO* p = dst + (unsigned)(u * a.ustride + j * a.stride);
      unsigned step = (unsigned)(R1 * a.stride);
      int c = 0;
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) {
        float2 val = y[k2];
        if (INV) val.y = -val.y;
        val = RAMPS ? cmul(val, s_post[j + R1 * k2]) : cscale(val, a.scale);
        SO::stnt(p + k2 * step, val);
        if constexpr (CHECK) c += stored_nonzero<SO>(val) ? 1 : 0;
      }
      if constexpr (CHECK) atomicAdd(&cnt[T + t], c);
    }
    __syncthreads();
    if constexpr (CHECK) {
      // this tile's counts are complete; the next tile counts into the other
      // parity, and this one is cleared before the tile after next uses it
      if (active && j == 0) {
        const int ci = cnt[t], co = cnt[T + t];
        if ((ci != 0 && ci != N) || (co != 0 && co != N)) atomicOr(a.flag, 1u);
        cnt[t] = 0;
        cnt[T + t] = 0;
      }
    }
  }
}

struct ZProjArgs {
  void* data;          // image-space iterate after the forward x and y passes (ST::T)
  const void* meas;    // measurements (k-space), same layout and storage
  int stride, ustride;
  long long wstride, vstride;
  int U, tiles, tiles_per_wg;
  const float2* pk;    // [v][nz] exp(+i phi_ksp,z)
  const float2* pi;    // [v][nz] exp(+i phi_img,z)
  float scale;         // 1/sqrt(nz)
  const float2* tw;
  double* partial;     // one partial sum of |x - y|^2 per workgroup
  long long cap;       // doubles available at partial (host-side check, not read by the kernel)
};

// Forward z pass -> hard projection -> adjoint z pass on tiles of z lines.
// The forward output ramp conj(pi) and the adjoint input ramp pi cancel on
// unsampled entries, so the projection compares x*scale with y*pi and, where
// sampled, substitutes y*pi (|pi| = 1, so |x*conj(pi) - y| = |x - y*pi|).
template <int R1, int R2, class ST, int MINT>
__global__ void __launch_bounds__((Layout<R1, R2, MINT>::THREADS)) k_fft_z_project(ZProjArgs a) {
  using L = Layout<R1, R2, MINT>;
  using E = typename ST::T;
  constexpr int N = L::N, T = L::T, TH = L::THREADS;
  __shared__ float2 lds[T * L::LS];
  __shared__ float2 s_tw[N], s_pk[N], s_pi[N];
  __shared__ double red[TH / 64];
  const int tid = threadIdx.x;
  const int t = tid % T;
  const int j = tid / T;
  const long long v = blockIdx.z;
  for (int i = tid; i < N; i += TH) {
    s_tw[i] = a.tw[i];
    s_pk[i] = a.pk[v * N + i];
    s_pi[i] = a.pi[v * N + i];
  }
  const long long row = v * a.vstride + (long long)blockIdx.y * a.wstride;
  E* __restrict__ data = (E*)a.data + row;
  const E* __restrict__ meas = (const E*)a.meas + row;
  const int tile0 = blockIdx.x * a.tiles_per_wg;
  const int ntiles = min(a.tiles_per_wg, a.tiles - tile0);
  double err = 0.0;
  // the next tile's iterate (stage-1 lanes) and measurements (stage-2 lanes:
  // k-space sample j + R1*k2) are loaded while the current tile computes
  float2 xn[R1];
  typename ST::T mn[R2];
  auto load = [&](int tile, float2* x, typename ST::T* m) {
    int u = tile * T + t;
    if (u < a.U && j < R2) {
      const E* p = data + (unsigned)(u * a.ustride + j * a.stride);
      unsigned step = (unsigned)(R2 * a.stride);
#pragma unroll
      for (int j1 = 0; j1 < R1; ++j1) x[j1] = ST::ldnt(p + j1 * step);
    }
    if (u < a.U && j < R1) {
      const E* mp = meas + (unsigned)(u * a.ustride + j * a.stride);
      unsigned step = (unsigned)(R1 * a.stride);
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) m[k2] = ST::rawnt(mp + k2 * step);
    }
  };
  load(tile0, xn, mn);
  __syncthreads();
  for (int it = 0; it < ntiles; ++it) {
    const int u = (tile0 + it) * T + t;
    const bool active = u < a.U;
    if (active && j < R2) {
      float2 x[R1];
#pragma unroll
      for (int j1 = 0; j1 < R1; ++j1) x[j1] = cmul(xn[j1], s_pk[j1 * R2 + j]);
      SmallDFT<R1>::run(x);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) lds[t * L::LS + k1 * L::R2P + j] = cmul(x[k1], s_tw[j * k1]);
    }
    typename ST::T mc[R2];
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) mc[k2] = mn[k2];
    __syncthreads();
    if (it + 1 < ntiles) load(tile0 + it + 1, xn, mn);
    float2 y[R2];
    if (active && j < R1) {
#pragma unroll
      for (int j2 = 0; j2 < R2; ++j2) y[j2] = lds[t * L::LS + j * L::R2P + j2];
      SmallDFT<R2>::run(y);
      // k-space sample k = j + R1*k2: project, then start the adjoint
      // (R2-point inverse DFT over k2 for the fixed k1 = j)
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) {
        float2 xv = cscale(y[k2], a.scale);
        float2 m = ST::ld(&mc[k2]);
        if (m.x != 0.0f || m.y != 0.0f) {
          float2 mpv = cmul(m, s_pi[j + R1 * k2]);
          float2 d = csub(xv, mpv);
          err += (double)cnorm(d);
          xv = mpv;
        }
        y[k2] = make_float2(xv.x, -xv.y);  // conjugated: inverse DFT via the forward kernel
      }
      SmallDFT<R2>::run(y);
    }
    __syncthreads();  // every stage-2 read of lds is done
    if (active && j < R1) {
#pragma unroll
      for (int j2 = 0; j2 < R2; ++j2)
        lds[t * L::LS + j * L::R2P + j2] = cmul(make_float2(y[j2].x, -y[j2].y), cconj(s_tw[j * j2]));
    }
    __syncthreads();
    if (active && j < R2) {
      float2 x[R1];
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) {
        float2 q = lds[t * L::LS + k1 * L::R2P + j];
        x[k1] = make_float2(q.x, -q.y);
      }
      SmallDFT<R1>::run(x);
      E* p = data + (unsigned)(u * a.ustride + j * a.stride);
      unsigned step = (unsigned)(R2 * a.stride);
#pragma unroll
      for (int j1 = 0; j1 < R1; ++j1) {
        int idx = j + R2 * j1;
        float2 r = make_float2(x[j1].x, -x[j1].y);
        ST::stnt(p + j1 * step, cscale(cmulc(r, s_pk[idx]), a.scale));
      }
    }
    __syncthreads();
  }
  // workgroup reduction of err (fixed order: deterministic)
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) err += __shfl_down(err, off, 64);
  if ((tid & 63) == 0) red[tid >> 6] = err;
  __syncthreads();
  if (tid == 0) {
    double s = 0.0;
    for (int w = 0; w < TH / 64; ++w) s += red[w];
    size_t bid = ((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    a.partial[bid] = s;
  }
}

// The same fused pass with R2 lanes per line instead of max(R1, R2), for
// R1 = q R2: stage 1 (R1-point DFTs, one per lane) keeps every lane busy and
// each lane takes q of the R1 R2-point DFTs of stage 2 (k1 = j + R2 u). With
// half the lanes per line, T = 32 lines make a 256-thread workgroup and three
// of them fit a CU: half storage then has the loads in flight that complex64
// storage gets from the wide layout (whose stage 1 leaves half the lanes idle).
template <int R1, int R2, class ST, int T>
__global__ void __launch_bounds__(T * R2) k_fft_z_project_nw(ZProjArgs a) {
  static_assert(R1 % R2 == 0 && R1 > R2, "narrow layout: R1 a multiple of R2");
  constexpr int N = R1 * R2, TH = T * R2, Q = R1 / R2;
  constexpr int R2P = R2 + 1, LS = R1 * R2P + 1;
  using E = typename ST::T;
  __shared__ float2 lds[T * LS];
  __shared__ float2 s_tw[N], s_pk[N], s_pi[N];
  __shared__ double red[(TH + 63) / 64];  // (a 32-lane workgroup is one partial wave)
  const int tid = threadIdx.x;
  const int t = tid % T;
  const int j = tid / T;  // 0 .. R2-1
  const long long v = blockIdx.z;
  for (int i = tid; i < N; i += TH) {
    s_tw[i] = a.tw[i];
    s_pk[i] = a.pk[v * N + i];
    s_pi[i] = a.pi[v * N + i];
  }
  const long long row = v * a.vstride + (long long)blockIdx.y * a.wstride;
  E* __restrict__ data = (E*)a.data + row;
  const E* __restrict__ meas = (const E*)a.meas + row;
  const int tile0 = blockIdx.x * a.tiles_per_wg;
  const int ntiles = min(a.tiles_per_wg, a.tiles - tile0);
  double err = 0.0;
  // next tile: iterate samples j + R2 j1 (stage 1) and measurements at
  // k = k1 + R1 k2 for this lane's stage-2 rows k1 = j + R2 u
  float2 xn[R1];
  E mn[Q][R2];
  // (the measurements of the next tile are loaded once stage 2 has used the
  // current ones: no second register copy of them)
  auto loadx = [&](int tile) {
    const int u = tile * T + t;
    if (u < a.U) {
      const E* p = data + (unsigned)(u * a.ustride + j * a.stride);
      const unsigned step = (unsigned)(R2 * a.stride);
#pragma unroll
      for (int j1 = 0; j1 < R1; ++j1) xn[j1] = ST::ldnt(p + j1 * step);
    }
  };
  auto loadm = [&](int tile) {
    const int u = tile * T + t;
    if (u < a.U) {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const E* mp = meas + (unsigned)(u * a.ustride + (j + R2 * q) * a.stride);
        const unsigned ms = (unsigned)(R1 * a.stride);
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) mn[q][k2] = ST::rawnt(mp + k2 * ms);
      }
    }
  };
  loadx(tile0);
  loadm(tile0);
  __syncthreads();
  for (int it = 0; it < ntiles; ++it) {
    const int u = (tile0 + it) * T + t;
    const bool active = u < a.U;
    if (active) {
      float2 x[R1];
#pragma unroll
      for (int j1 = 0; j1 < R1; ++j1) x[j1] = cmul(xn[j1], s_pk[j1 * R2 + j]);
      SmallDFT<R1>::run(x);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) lds[t * LS + k1 * R2P + j] = cmul(x[k1], s_tw[j * k1]);
    }
    __syncthreads();
    if (it + 1 < ntiles) loadx(tile0 + it + 1);
    if (active) {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int k1 = j + R2 * q;
        float2 y[R2];
#pragma unroll
        for (int j2 = 0; j2 < R2; ++j2) y[j2] = lds[t * LS + k1 * R2P + j2];
        SmallDFT<R2>::run(y);
        // k-space sample k = k1 + R1 k2: project, then the adjoint's R2-point DFT
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) {
          float2 xv = cscale(y[k2], a.scale);
          const float2 m = ST::ld(&mn[q][k2]);
          if (m.x != 0.0f || m.y != 0.0f) {
            const float2 mpv = cmul(m, s_pi[k1 + R1 * k2]);
            const float2 d = csub(xv, mpv);
            err += (double)cnorm(d);
            xv = mpv;
          }
          y[k2] = make_float2(xv.x, -xv.y);  // conjugated: inverse DFT via the forward kernel
        }
        SmallDFT<R2>::run(y);
        // each lane rewrites only the rows k1 it read (no other lane reads them)
#pragma unroll
        for (int j2 = 0; j2 < R2; ++j2)
          lds[t * LS + k1 * R2P + j2] = cmul(make_float2(y[j2].x, -y[j2].y), cconj(s_tw[k1 * j2]));
      }
    }
    if (it + 1 < ntiles) loadm(tile0 + it + 1);
    __syncthreads();
    if (active) {
      float2 x[R1];
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) {
        const float2 q = lds[t * LS + k1 * R2P + j];
        x[k1] = make_float2(q.x, -q.y);
      }
      SmallDFT<R1>::run(x);
      E* p = data + (unsigned)(u * a.ustride + j * a.stride);
      const unsigned step = (unsigned)(R2 * a.stride);
#pragma unroll
      for (int j1 = 0; j1 < R1; ++j1) {
        const int idx = j + R2 * j1;
        const float2 r = make_float2(x[j1].x, -x[j1].y);
        ST::stnt(p + j1 * step, cscale(cmulc(r, s_pk[idx]), a.scale));
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) err += __shfl_down(err, off, 64);
  if ((tid & 63) == 0) red[tid >> 6] = err;
  __syncthreads();
  if (tid == 0) {
    double s = 0.0;
    for (int w = 0; w < (TH + 63) / 64; ++w) s += red[w];
    const size_t bid = ((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    a.partial[bid] = s;
  }
}

// Adjoint y pass fused with the subspace projection C = X S (k_project's
// C[k][n] = sum_q X[q][n] S[q][k]): the adjoint image is not stored, only its
// K coefficient planes. A workgroup owns YP_T x-neighbour y-lines of one z
// slice and walks the volumes VG at a time: each group of T*P threads runs
// one volume's inverse line FFTs exactly as k_fft_pass does (same stage order,
// twiddles and ramps: the values are those the unfused pass would store),
// leaving the outputs in LDS, and then every thread adds the VG volumes'
// contributions to the K coefficients of its VPT voxels, held in registers
// until the last volume. The input stays where it is (the z pass's output).
// VG = 2: 512-thread workgroups, 8 voxels (112 accumulators) per thread, one
// workgroup per CU (80 KB of LDS at ny = 256). At 256x256x128x67 this takes
// 1.41 / 1.60 ms per launch for half / complex64 storage with the parity-sign
// input ramps (1.51 / 1.69 ms multiplying by the table); VG = 4 (1024
// threads, 56 accumulators, spilling at the 128 registers such a workgroup
// allows) took 1.75 / 2.83 ms.
constexpr int YP_T = 16, YP_VG = 2;

struct YProjArgs {
  const void* src;     // (x, ky, z) data of every volume (ST::T), volume pitch vstride
  float2* C;           // K planes of N voxels, x fastest
  const float2* pre;   // [v][ny] input ramps (used as given)
  const float2* post;  // [v][ny] output ramps (conjugated, times scale)
  const float2* tw;    // exp(-2 pi i m / ny)
  long long nxy, vstride, N;
  int nx, nz, K, nq;
  float scale;
  int tiles_x;         // ceil(nx / YP_T) x tiles per z slice
  int per_xcd;         // work items per XCD (grid = 8 per_xcd)
};

template <int R1, int R2, int VG_>
struct YProjShape {
  static constexpr int N = R1 * R2, P = R1 > R2 ? R1 : R2, T = YP_T, VG = VG_;
  static constexpr int GT = T * P, TH = VG * GT;
  static constexpr int R2P = R2 + 1, LS = R1 * R2P + 1;
  static constexpr int VPT = T * N / TH;  // voxels accumulated per thread
  static constexpr int NR = 2 * VG * N;   // ramp entries staged per step (input and output ramps)
};

// SGN: the input ramps are the parity pattern (-1)^(y + N/2). Stage 1's inputs
// y = j1 R2 + j (R2 even) then share lane j's sign, which the linear stage-1
// DFT carries to its outputs: it is folded into that lane's twiddles
// (s_tw[j R1 + k1] = sign(j) W^-(j k1)) and only the output ramps are staged.
template <int R1, int R2, int VG_, class ST, int KK, bool RS, bool SGN>
__global__ void __launch_bounds__((YProjShape<R1, R2, VG_>::TH)) k_fft_y_adjoint_project(YProjArgs a,
                                                                                        const float2* __restrict__ S) {
  using L = YProjShape<R1, R2, VG_>;
  constexpr int N = L::N, T = L::T, VG = L::VG, GT = L::GT, TH = L::TH, R2P = L::R2P, LS = L::LS, VPT = L::VPT;
  static_assert(VPT * TH == T * N, "every thread accumulates the same number of voxels");
  static_assert(!SGN || R2 % 2 == 0, "the parity of y = j1 R2 + j is that of j");
  constexpr int NR = SGN ? VG * N : L::NR, RPT = (NR + TH - 1) / TH;
  using E = typename ST::T;
  __shared__ float2 lds[VG][T * LS];
  __shared__ float2 s_tw[N];
  __shared__ float2 s_rp[2][VG][N];  // [pre, post][group][index] (SGN: post only, in s_rp[0])
  const int tid = threadIdx.x;
  const int g = tid / GT;
  const int t = (tid % GT) % T, j = (tid % GT) / T;
  // XCD-contiguous work items: the x tiles of one z slice run side by side on one XCD
  const int w = (int)(blockIdx.x % 8) * a.per_xcd + (int)(blockIdx.x / 8);
  if (w >= a.tiles_x * a.nz) return;
  const int x0 = (w % a.tiles_x) * T;
  const long long z = w / a.tiles_x;
  for (int i = tid; i < N; i += TH) {
    if (SGN) {
      const int jj = i / R1, k1 = i % R1;  // (jj < R2)
      const float2 w = cconj(a.tw[jj * k1]);
      s_tw[i] = ((jj + N / 2) & 1) ? make_float2(-w.x, -w.y) : w;
    } else {
      s_tw[i] = cconj(a.tw[i]);
    }
  }
  const bool line = x0 + t < a.nx;
  const E* __restrict__ src = (const E*)a.src + z * a.nxy + x0 + t + (long long)j * a.nx;
  const long long step = (long long)R2 * a.nx;
  E xr[R1];
  float2 rr[RPT];
  auto load = [&](int v0) {
    const int v = v0 + g;
    if (line && v < a.nq && j < R2) {
      const E* p = src + (long long)v * a.vstride;
#pragma unroll
      for (int j1 = 0; j1 < R1; ++j1) xr[j1] = ST::rawnt(p + j1 * step);
    }
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int e = tid + TH * q;
      if (e < NR) {
        const int tb = SGN ? 1 : e / (VG * N), gg = (e / N) % VG, i = e % N;
        const long long vv = min(v0 + gg, a.nq - 1);
        rr[q] = tb == 0 ? a.pre[vv * N + i] : cscale(cconj(a.post[vv * N + i]), a.scale);
      }
    }
  };
  constexpr int PO = SGN ? 0 : 1;  // s_rp slot of the output ramps
  float2 acc[VPT][KK];
#pragma unroll
  for (int i = 0; i < VPT; ++i)
#pragma unroll
    for (int k = 0; k < KK; ++k) acc[i][k] = make_float2(0.f, 0.f);
  load(0);
  for (int v0 = 0; v0 < a.nq; v0 += VG) {
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int e = tid + TH * q;
      if (e < NR) (&s_rp[0][0][0])[e] = rr[q];
    }
    // tables ready; the previous step's readers of lds are done
    __syncthreads();
    const bool vok = v0 + g < a.nq;
    if (line && vok && j < R2) {
      float2 x[R1];
#pragma unroll
      for (int j1 = 0; j1 < R1; ++j1) {
        x[j1] = SGN ? ST::cvt(xr[j1]) : cmul(ST::cvt(xr[j1]), s_rp[0][g][j1 * R2 + j]);
        x[j1].y = -x[j1].y;
      }
      SmallDFT<R1>::run(x);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) {
        float2 y = x[k1];
        y.y = -y.y;
        lds[g][t * LS + k1 * R2P + j] = cmul(y, s_tw[SGN ? j * R1 + k1 : j * k1]);
      }
    }
    if (v0 + VG < a.nq) load(v0 + VG);
    __syncthreads();
    // stage 2 in place: lane j reads row j and writes its outputs y = j + R1 k2 back into it
    if (line && vok && j < R1) {
      float2 y[R2];
#pragma unroll
      for (int j2 = 0; j2 < R2; ++j2) {
        y[j2] = lds[g][t * LS + j * R2P + j2];
        y[j2].y = -y[j2].y;
      }
      SmallDFT<R2>::run(y);
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) {
        float2 val = y[k2];
        val.y = -val.y;
        lds[g][t * LS + j * R2P + k2] = cmul(val, s_rp[PO][g][j + R1 * k2]);
      }
    }
    __syncthreads();
    // one volume at a time (a rolled loop: the LDS reads of every volume are
    // not hoisted ahead of the FMAs, which would not fit the registers)
#pragma unroll 1
    for (int gg = 0; gg < VG; ++gg) {
      const int v = v0 + gg;
      if (v >= a.nq) break;
      float2 sv[KK];
#pragma unroll
      for (int k = 0; k < KK; ++k) sv[k] = S[v + (long long)min(k, a.K - 1) * a.nq];
#pragma unroll
      for (int i = 0; i < VPT; ++i) {
        const int e = tid + TH * i, tt = e % T, yy = e / T;
        if (x0 + tt >= a.nx) continue;
        const float2 xv = lds[gg][tt * LS + (yy % R1) * R2P + yy / R1];
#pragma unroll
        for (int k = 0; k < KK; ++k) {
          acc[i][k].x = fmaf(xv.x, sv[k].x, acc[i][k].x);
          acc[i][k].y = fmaf(xv.y, sv[k].x, acc[i][k].y);
          if (!RS) {
            acc[i][k].x = fmaf(-xv.y, sv[k].y, acc[i][k].x);
            acc[i][k].y = fmaf(xv.x, sv[k].y, acc[i][k].y);
          }
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < VPT; ++i) {
    const int e = tid + TH * i, tt = e % T, yy = e / T;
    if (x0 + tt >= a.nx) continue;
    float2* c = a.C + z * a.nxy + (long long)yy * a.nx + x0 + tt;
#pragma unroll
    for (int k = 0; k < KK; ++k)
      if (k < a.K) c[(long long)k * a.N] = acc[i][k];
  }
}

template <int R1, int R2, int VG, class ST, bool SGN>
hipError_t launch_yproj_st(YProjArgs a, const float2* S, bool rs, hipStream_t s) {
  using L = YProjShape<R1, R2, VG>;
  a.tiles_x = (a.nx + L::T - 1) / L::T;
  const long long W = (long long)a.tiles_x * a.nz;
  a.per_xcd = (int)((W + 7) / 8);
  const dim3 grid((unsigned)(8 * a.per_xcd));
  if (a.K == 7 && rs)
    hipLaunchKernelGGL((k_fft_y_adjoint_project<R1, R2, VG, ST, 7, true, SGN>), grid, dim3(L::TH), 0, s, a, S);
  else if (a.K == 7)
    hipLaunchKernelGGL((k_fft_y_adjoint_project<R1, R2, VG, ST, 7, false, SGN>), grid, dim3(L::TH), 0, s, a, S);
  else
    hipLaunchKernelGGL((k_fft_y_adjoint_project<R1, R2, VG, ST, 8, false, SGN>), grid, dim3(L::TH), 0, s, a, S);
  return hipGetLastError();
}

template <int R1, int R2>
hipError_t launch_yproj(YProjArgs a, const float2* S, bool rs, bool h16, bool sgn, hipStream_t s) {
  if (sgn)
    return h16 ? launch_yproj_st<R1, R2, YP_VG, StoreF16, true>(a, S, rs, s)
               : launch_yproj_st<R1, R2, YP_VG, StoreF32, true>(a, S, rs, s);
  return h16 ? launch_yproj_st<R1, R2, YP_VG, StoreF16, false>(a, S, rs, s)
             : launch_yproj_st<R1, R2, YP_VG, StoreF32, false>(a, S, rs, s);
}

template <int R1, int R2, bool TF, bool IV, bool RP, class ST, int MINT>
hipError_t launch_one(PassArgs a, long long W, long long nvols, hipStream_t s) {
  using L = Layout<R1, R2, MINT>;
  a.tiles = (a.U + L::T - 1) / L::T;
  // tiles per workgroup (tunable): 3 for half storage's 32-line tiles, 2 for
  // complex64 (256-point y passes at 256x256x128x67, swept over 1-8: half
  // storage 5.72 TB/s at 3 against 5.29 at 2 and 4; complex64 5.41 at 2)
  const int want = a.tiles_per_wg > 0 ? a.tiles_per_wg : (MINT > 1 ? 3 : 2);
  a.tiles_per_wg = a.tiles >= want ? want : a.tiles;
  dim3 grid((a.tiles + a.tiles_per_wg - 1) / a.tiles_per_wg, (unsigned)W, (unsigned)nvols);
  hipLaunchKernelGGL((k_fft_pass<R1, R2, TF, IV, RP, ST, MINT>), grid, dim3(L::THREADS), 0, s, a);
  return hipGetLastError();
}

// x-line checked inverse pass from complex64 (into complex64 or half storage)
template <int R1, int R2>
hipError_t launch_checked(PassArgs a, long long W, long long nvols, bool h16_out, hipStream_t s) {
  using L = Layout<R1, R2, 1>;
  a.tiles = (a.U + L::T - 1) / L::T;
  const int want = a.tiles_per_wg > 0 ? a.tiles_per_wg : 4;
  a.tiles_per_wg = a.tiles >= want ? want : a.tiles;
  dim3 grid((a.tiles + a.tiles_per_wg - 1) / a.tiles_per_wg, (unsigned)W, (unsigned)nvols);
  if (h16_out)
    hipLaunchKernelGGL((k_fft_pass<R1, R2, false, true, true, StoreF32, 1, StoreF16, true>), grid, dim3(L::THREADS), 0,
                       s, a);
  else
    hipLaunchKernelGGL((k_fft_pass<R1, R2, false, true, true, StoreF32, 1, StoreF32, true>), grid, dim3(L::THREADS), 0,
                       s, a);
  return hipGetLastError();
}

template <int R1, int R2>
hipError_t launch_pass(PassArgs a, long long W, long long nvols, bool tfast, bool inv, bool h16, hipStream_t s) {
  const bool ramps = a.pre != nullptr;
  if (h16) {
    // half storage exists for the signal model's ramped passes only; the
    // strided (tfast) passes take 32 x-neighbour lines per workgroup
    if (!ramps) return hipErrorInvalidValue;
    if (tfast) return inv ? launch_one<R1, R2, true, true, true, StoreF16, 32>(a, W, nvols, s)
                          : launch_one<R1, R2, true, false, true, StoreF16, 32>(a, W, nvols, s);
    return inv ? launch_one<R1, R2, false, true, true, StoreF16, 1>(a, W, nvols, s)
               : launch_one<R1, R2, false, false, true, StoreF16, 1>(a, W, nvols, s);
  }
#define FFT_LAUNCH(TF, IV, RP) return launch_one<R1, R2, TF, IV, RP, StoreF32, 1>(a, W, nvols, s)
  if (tfast) {
    if (inv) { if (ramps) FFT_LAUNCH(true, true, true); else FFT_LAUNCH(true, true, false); }
    else { if (ramps) FFT_LAUNCH(true, false, true); else FFT_LAUNCH(true, false, false); }
  } else {
    if (inv) { if (ramps) FFT_LAUNCH(false, true, true); else FFT_LAUNCH(false, true, false); }
    else { if (ramps) FFT_LAUNCH(false, false, true); else FFT_LAUNCH(false, false, false); }
  }
#undef FFT_LAUNCH
}

template <int R1, int R2, class ST, int MINT>
hipError_t launch_zproj_as(ZProjArgs a, long long W, long long nvols, hipStream_t s, size_t* nparts) {
  using L = Layout<R1, R2, MINT>;
  a.tiles = (a.U + L::T - 1) / L::T;
  // size query: one tile per workgroup gives the most workgroups of any setting;
  // 6 by default (complex64 at 256x256x128x67, swept over 1-8: 5.57 TB/s
  // against 5.42 at 4)
  const int want = a.partial == nullptr ? 1 : a.tiles_per_wg > 0 ? a.tiles_per_wg : 6;
  a.tiles_per_wg = a.tiles >= want ? want : a.tiles;
  dim3 grid((a.tiles + a.tiles_per_wg - 1) / a.tiles_per_wg, (unsigned)W, (unsigned)nvols);
  *nparts = (size_t)grid.x * grid.y * grid.z;
  if (a.partial == nullptr) return hipSuccess;
  if ((long long)*nparts > a.cap) return hipErrorInvalidValue;  // partial sums would not fit
  hipLaunchKernelGGL((k_fft_z_project<R1, R2, ST, MINT>), grid, dim3(L::THREADS), 0, s, a);
  return hipGetLastError();
}

// narrow-layout launch: 32 lines per workgroup, 3 tiles per workgroup by
// default (at 256x256x128x67 in half storage: 1.17 ms per launch, against 1.25
// at 2 or 4 tiles and 1.66 ms for the wide layout; 64 lines per 512-thread
// workgroup takes 1.33-1.39 ms, its 139 registers per lane allowing one
// workgroup per CU, or spilling when held to 128)
template <int R1, int R2, class ST>
hipError_t launch_zproj_nw(ZProjArgs a, long long W, long long nvols, hipStream_t s, size_t* nparts) {
  constexpr int T = 32;
  a.tiles = (a.U + T - 1) / T;
  const int want = a.tiles_per_wg > 0 ? a.tiles_per_wg : 3;
  a.tiles_per_wg = a.tiles >= want ? want : a.tiles;
  dim3 grid((a.tiles + a.tiles_per_wg - 1) / a.tiles_per_wg, (unsigned)W, (unsigned)nvols);
  *nparts = (size_t)grid.x * grid.y * grid.z;
  if ((long long)*nparts > a.cap) return hipErrorInvalidValue;  // partial sums would not fit
  hipLaunchKernelGGL((k_fft_z_project_nw<R1, R2, ST, T>), grid, dim3(T * R2), 0, s, a);
  return hipGetLastError();
}

template <int R1, int R2>
hipError_t launch_zproj(ZProjArgs a, long long W, long long nvols, bool h16, hipStream_t s, size_t* nparts) {
  // half storage with R1 a multiple of R2 takes the narrow layout; the size
  // query stays with the wide layout, whose workgroup count bounds both
  if constexpr (R1 > R2 && R1 % R2 == 0 && R2 >= 2 && R2 <= 8) {  // whole waves: 32 R2 >= 64 threads
    if (h16 && a.partial != nullptr) return launch_zproj_nw<R1, R2, StoreF16>(a, W, nvols, s, nparts);
  }
  return h16 ? launch_zproj_as<R1, R2, StoreF16, 32>(a, W, nvols, s, nparts)
             : launch_zproj_as<R1, R2, StoreF32, 1>(a, W, nvols, s, nparts);
}

}  // namespace

namespace llr {

bool fft_size_supported(size_t n) {
  switch (n) {
    case 2: case 3: case 4: case 6: case 8: case 12: case 16: case 24: case 32: case 48: case 64: case 96:
    case 128: case 192: case 256: case 384: case 512:
      return true;
    default:
      return false;
  }
}

// Dispatch an axis pass for line length n.
static hipError_t dispatch_pass(int n, const PassArgs& a, long long W, long long nv, bool tfast, bool inv, bool h16,
                                hipStream_t s) {
  switch (n) {
    case 2: return launch_pass<2, 1>(a, W, nv, tfast, inv, h16, s);
    case 4: return launch_pass<2, 2>(a, W, nv, tfast, inv, h16, s);
    case 8: return launch_pass<4, 2>(a, W, nv, tfast, inv, h16, s);
    case 16: return launch_pass<4, 4>(a, W, nv, tfast, inv, h16, s);
    case 32: return launch_pass<8, 4>(a, W, nv, tfast, inv, h16, s);
    case 64: return launch_pass<8, 8>(a, W, nv, tfast, inv, h16, s);
    case 128: return launch_pass<16, 8>(a, W, nv, tfast, inv, h16, s);
    case 256: return launch_pass<16, 16>(a, W, nv, tfast, inv, h16, s);
    case 512: return launch_pass<32, 16>(a, W, nv, tfast, inv, h16, s);
    case 3: return launch_pass<3, 1>(a, W, nv, tfast, inv, h16, s);
    case 6: return launch_pass<3, 2>(a, W, nv, tfast, inv, h16, s);
    case 12: return launch_pass<6, 2>(a, W, nv, tfast, inv, h16, s);
    case 24: return launch_pass<6, 4>(a, W, nv, tfast, inv, h16, s);
    case 48: return launch_pass<12, 4>(a, W, nv, tfast, inv, h16, s);
    case 96: return launch_pass<12, 8>(a, W, nv, tfast, inv, h16, s);
    case 192: return launch_pass<12, 16>(a, W, nv, tfast, inv, h16, s);
    case 384: return launch_pass<12, 32>(a, W, nv, tfast, inv, h16, s);
    default: return hipErrorInvalidValue;
  }
}

static hipError_t dispatch_checked(int n, const PassArgs& a, long long W, long long nv, bool h16_out, hipStream_t s) {
  switch (n) {
    case 2: return launch_checked<2, 1>(a, W, nv, h16_out, s);
    case 4: return launch_checked<2, 2>(a, W, nv, h16_out, s);
    case 8: return launch_checked<4, 2>(a, W, nv, h16_out, s);
    case 16: return launch_checked<4, 4>(a, W, nv, h16_out, s);
    case 32: return launch_checked<8, 4>(a, W, nv, h16_out, s);
    case 64: return launch_checked<8, 8>(a, W, nv, h16_out, s);
    case 128: return launch_checked<16, 8>(a, W, nv, h16_out, s);
    case 256: return launch_checked<16, 16>(a, W, nv, h16_out, s);
    case 512: return launch_checked<32, 16>(a, W, nv, h16_out, s);
    case 3: return launch_checked<3, 1>(a, W, nv, h16_out, s);
    case 6: return launch_checked<3, 2>(a, W, nv, h16_out, s);
    case 12: return launch_checked<6, 2>(a, W, nv, h16_out, s);
    case 24: return launch_checked<6, 4>(a, W, nv, h16_out, s);
    case 48: return launch_checked<12, 4>(a, W, nv, h16_out, s);
    case 96: return launch_checked<12, 8>(a, W, nv, h16_out, s);
    case 192: return launch_checked<12, 16>(a, W, nv, h16_out, s);
    case 384: return launch_checked<12, 32>(a, W, nv, h16_out, s);
    default: return hipErrorInvalidValue;
  }
}

// One pass along `axis` of nvols volumes of size dims; src may equal dst.
// pre and post must both be given (ramped signal-model pass) or both be null.
llr_status fft_axis_pass(llr_ctx* ctx, const void* src, void* dst, const size_t* dims, size_t nvols, int axis,
                         bool inverse, const float2* pre, bool pre_conj, const float2* post, bool post_conj,
                         float scale, bool h16, const float2* twiddles) {
  int n = (int)dims[axis];
  if (!fft_size_supported((size_t)n))
    return fail(LLR_ERR_UNSUPPORTED, "FFT length " + std::to_string(n) + " is not supported (2^a * 3^b with a radix of at most 32, up to 512)");
  if ((pre == nullptr) != (post == nullptr)) return fail(LLR_ERR_ARG, "ramp tables must be given together");
  if (h16 && pre == nullptr) return fail(LLR_ERR_UNSUPPORTED, "half storage is for the signal model's passes only");
  if ((double)dims[0] * dims[1] * dims[2] >= 4294967296.0)
    return fail(LLR_ERR_UNSUPPORTED, "volume larger than 2^32 voxels");
  const float2* tw = twiddles;
  if (!tw) {
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
  long long W;
  bool tfast;
  if (axis == 0) {
    a.stride = 1;
    a.U = (int)ny;
    a.ustride = (int)nx;
    W = nz;
    a.wstride = nx * ny;
    tfast = false;
  } else if (axis == 1) {
    a.stride = (int)nx;
    a.U = (int)nx;
    a.ustride = 1;
    W = nz;
    a.wstride = nx * ny;
    tfast = true;
  } else {
    a.stride = (int)(nx * ny);
    a.U = (int)nx;
    a.ustride = 1;
    W = ny;
    a.wstride = nx;
    tfast = true;
  }
  a.tiles_per_wg = ctx->fft_tpw[axis];
  hipError_t e = dispatch_pass(n, a, W, (long long)nvols, tfast, inverse, h16, ctx->stream);
  if (e != hipSuccess) return hip_fail(e, "fft axis pass launch");
  return LLR_OK;
}

llr_status fft_x_inverse_checked(llr_ctx* ctx, const float2* src, void* dst, bool h16_out, const size_t* dims,
                                 size_t nvols, const float2* pre, const float2* post, float scale, unsigned* flag) {
  const int n = (int)dims[0];
  if (!fft_size_supported((size_t)n))
    return fail(LLR_ERR_UNSUPPORTED, "FFT length " + std::to_string(n) + " is not supported");
  if (!pre || !post || !flag) return fail(LLR_ERR_ARG, "checked x pass needs ramp tables and a flag");
  if ((double)dims[0] * dims[1] * dims[2] >= 4294967296.0)
    return fail(LLR_ERR_UNSUPPORTED, "volume larger than 2^32 voxels");
  const float2* tw = nullptr;
  llr_status st = get_twiddles(ctx, n, &tw);
  if (st) return st;
  const long long nx = dims[0], ny = dims[1], nz = dims[2];
  PassArgs a{};
  a.src = src;
  a.dst = dst;
  a.vstride = nx * ny * nz;
  a.pre = pre;
  a.post = post;
  a.pre_conj = 0;
  a.post_conj = 1;
  a.scale = scale;
  a.tw = tw;
  a.stride = 1;
  a.U = (int)ny;
  a.ustride = (int)nx;
  a.wstride = nx * ny;
  a.tiles_per_wg = ctx->fft_tpw[0];
  a.flag = flag;
  hipError_t e = dispatch_checked(n, a, nz, (long long)nvols, h16_out, ctx->stream);
  if (e != hipSuccess) return hip_fail(e, "checked x pass launch");
  return LLR_OK;
}

// Fused forward-z / hard projection / adjoint-z. Writes one partial error sum
// per workgroup into `partial` (size returned in nparts; partial == null only
// queries the size).
llr_status fft_zproj(llr_ctx* ctx, void* data, const void* meas, const size_t* dims, size_t nvols,
                     const float2* pk, const float2* pi, double* partial, size_t cap, size_t* nparts, bool h16) {
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
  a.stride = (int)(nx * ny);
  a.U = (int)nx;
  a.ustride = 1;
  a.wstride = nx;
  a.vstride = nx * ny * nz;
  a.pk = pk;
  a.pi = pi;
  a.scale = (float)(1.0 / std::sqrt((double)nz));
  a.tw = tw;
  a.partial = partial;
  a.cap = (long long)cap;
  a.tiles_per_wg = ctx->fft_tpw[2];
  hipError_t e;
  switch (n) {
    case 2: e = launch_zproj<2, 1>(a, ny, nvols, h16, ctx->stream, nparts); break;
    case 4: e = launch_zproj<2, 2>(a, ny, nvols, h16, ctx->stream, nparts); break;
    case 8: e = launch_zproj<4, 2>(a, ny, nvols, h16, ctx->stream, nparts); break;
    case 16: e = launch_zproj<4, 4>(a, ny, nvols, h16, ctx->stream, nparts); break;
    case 32: e = launch_zproj<8, 4>(a, ny, nvols, h16, ctx->stream, nparts); break;
    case 64: e = launch_zproj<8, 8>(a, ny, nvols, h16, ctx->stream, nparts); break;
    case 128: e = launch_zproj<16, 8>(a, ny, nvols, h16, ctx->stream, nparts); break;
    case 256: e = launch_zproj<16, 16>(a, ny, nvols, h16, ctx->stream, nparts); break;
    case 512: e = launch_zproj<32, 16>(a, ny, nvols, h16, ctx->stream, nparts); break;
    case 3: e = launch_zproj<3, 1>(a, ny, nvols, h16, ctx->stream, nparts); break;
    case 6: e = launch_zproj<3, 2>(a, ny, nvols, h16, ctx->stream, nparts); break;
    case 12: e = launch_zproj<6, 2>(a, ny, nvols, h16, ctx->stream, nparts); break;
    case 24: e = launch_zproj<6, 4>(a, ny, nvols, h16, ctx->stream, nparts); break;
    case 48: e = launch_zproj<12, 4>(a, ny, nvols, h16, ctx->stream, nparts); break;
    case 96: e = launch_zproj<12, 8>(a, ny, nvols, h16, ctx->stream, nparts); break;
    case 192: e = launch_zproj<12, 16>(a, ny, nvols, h16, ctx->stream, nparts); break;
    case 384: e = launch_zproj<12, 32>(a, ny, nvols, h16, ctx->stream, nparts); break;
    default: return fail(LLR_ERR_UNSUPPORTED, "z length " + std::to_string(n) + " is not a supported FFT length");
  }
  if (e == hipErrorInvalidValue && partial && *nparts > cap)
    return fail(LLR_ERR_ARG, "fused z pass needs " + std::to_string(*nparts) + " partial sums, buffer holds " +
                                 std::to_string(cap));
  if (e != hipSuccess) return hip_fail(e, "fft z projection launch");
  return LLR_OK;
}

bool fft_yproj_supported(size_t ny, int K) {
  if (K < 1 || K > 8) return false;
  switch (ny) {
    case 16: case 24: case 32: case 48: case 64: case 96: case 128: case 192: case 256:
      return true;
    default:
      return false;
  }
}

// Inverse y pass (input ramps pre, output ramps conj(post) * scale, as
// fft_axis_pass with pre_conj = 0, post_conj = 1) of every volume, projected:
// C[k][n] = sum_q X[q][n] S[q][k] of the pass's output X, which is not stored.
llr_status fft_y_adjoint_project(llr_ctx* ctx, const void* src, const size_t* dims, size_t nvols, const float2* pre,
                                 const float2* post, float scale, float2* C, const float2* S, int K, bool s_real,
                                 bool h16, bool pre_sign) {
  if (!fft_yproj_supported(dims[1], K)) return fail(LLR_ERR_UNSUPPORTED, "fused adjoint projection: y length or rank");
  if (!pre || !post || !C || !S) return fail(LLR_ERR_ARG, "fused adjoint projection needs ramps, planes and S");
  if ((double)dims[0] * dims[1] * dims[2] >= 4294967296.0)
    return fail(LLR_ERR_UNSUPPORTED, "volume larger than 2^32 voxels");
  const int n = (int)dims[1];
  const float2* tw = nullptr;
  llr_status st = get_twiddles(ctx, n, &tw);
  if (st) return st;
  YProjArgs a{};
  a.src = src;
  a.C = C;
  a.pre = pre;
  a.post = post;
  a.tw = tw;
  a.nxy = (long long)(dims[0] * dims[1]);
  a.vstride = a.nxy * (long long)dims[2];
  a.N = a.vstride;
  a.nx = (int)dims[0];
  a.nz = (int)dims[2];
  a.K = K;
  a.nq = (int)nvols;
  a.scale = scale;
  hipError_t e;
  switch (n) {
    case 16: e = launch_yproj<4, 4>(a, S, s_real, h16, pre_sign, ctx->stream); break;
    case 24: e = launch_yproj<6, 4>(a, S, s_real, h16, pre_sign, ctx->stream); break;
    case 32: e = launch_yproj<8, 4>(a, S, s_real, h16, pre_sign, ctx->stream); break;
    case 48: e = launch_yproj<12, 4>(a, S, s_real, h16, pre_sign, ctx->stream); break;
    case 64: e = launch_yproj<8, 8>(a, S, s_real, h16, pre_sign, ctx->stream); break;
    case 96: e = launch_yproj<12, 8>(a, S, s_real, h16, pre_sign, ctx->stream); break;
    case 128: e = launch_yproj<16, 8>(a, S, s_real, h16, pre_sign, ctx->stream); break;
    case 192: e = launch_yproj<12, 16>(a, S, s_real, h16, pre_sign, ctx->stream); break;
    default: e = launch_yproj<16, 16>(a, S, s_real, h16, pre_sign, ctx->stream); break;
  }
  if (e != hipSuccess) return hip_fail(e, "fused adjoint projection launch");
  return LLR_OK;
}

}  // namespace llr
