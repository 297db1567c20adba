// Multi-level 3-D Daubechies-2 wavelet soft threshold on gfx950 (the
// wavelet_denoise step of lib.rs:430-465).
//
// Filters and boundary rules follow the reference's CUDA kernels: analysis
// a[c] = sum_k lo_d[k] x[sym(2c+1-k)] with symmetric (half-sample) extension
// (kernels.cu:97-105, 185-203), nc = (n+3)/2 coefficients per band
// (kernels.cu:150-152), approximation in [0,nc) and detail in [nc,2nc);
// synthesis keeps the samples from F-2 on (kernels.cu:230-261) and is trimmed
// to the finer level's length (pywt waverecn). Every detail subband and the
// deepest approximation are soft-thresholded (shrink_complex, lib.rs:460-465).
//
// One level takes three passes over its data instead of one pass per axis and
// direction:
//   1. k_wav_xy_an: an LDS tile (with its 2-sample halo) of one z-slice is
//      filtered along x then y, giving the four xy-bands of the level, stored
//      with z untouched (band buffer B, 2ncx x 2ncy x nz).
//   2. k_wav_z: a tile of z-lines of B is analysed along z, shrunk and
//      synthesised back in place while it sits in LDS. For the lines of the LL
//      band the z-approximation (LLL) is the next level's input; it is written
//      out and the line keeps only the synthesis of its shrunk detail, which
//      is valid because synthesis is linear: syn(a', d') = syn(a', 0) + syn(0, d').
//   3. k_wav_xy_syn: the xy-bands are synthesised along y then x from an LDS
//      tile; the LL band picks up syn_z(a', 0) of the reconstructed deeper level
//      while it is loaded.
// Levels run coarse-to-fine on the way back; the deepest level handles its
// LLL band completely in pass 2.
#include <hip/hip_runtime.h>

#include <array>
#include <cmath>

#include "common.h"
#include "internal.h"

namespace {

struct Db2F {
  float dlo[4], dhi[4], rlo[4], rhi[4];
};

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

// symmetric half-sample extension; always returns an index in [0, n)
__device__ __forceinline__ int sym(int i, int n) {
  if (i < 0) i = -i - 1;
  if (i >= n) i = 2 * n - i - 1;
  if (i < 0) i = 0;
  return i;
}

// x * max(|x| - alpha, 0) / |x| = x * max(1 - alpha / |x|, 0), with 1/|x| from
// the hardware reciprocal square root (about 1 ulp); |x| = 0 gives rsq = inf,
// 1 - inf -> factor 0, so zero maps to zero (shrink_complex, lib.rs:460-465,
// returns NaN there).
__device__ __forceinline__ float2 shrink(float2 x, float alpha) {
  const float m2 = x.x * x.x + x.y * x.y;
  const float s = fmaxf(1.f - alpha * __builtin_amdgcn_rsqf(m2), 0.f);
  return make_float2(x.x * s, x.y * s);
}

// x <- shrink(x) for every stored value (a level-0 transform: the volume is its
// own approximation)
template <class ST>
__global__ void __launch_bounds__(256) k_shrink_store(typename ST::T* x, long long n, float alpha) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    ST::st(x + i, shrink(ST::ld(x + i), alpha));
}

__device__ __forceinline__ void fma2(float2& acc, float w, float2 v) {
  acc.x += w * v.x;
  acc.y += w * v.y;
}

constexpr int AN_TCX = 32, AN_TCY = 16;              // coefficient tile of pass 1
constexpr int AN_W = 2 * AN_TCX + 2, AN_H = 2 * AN_TCY + 2;
constexpr int SY_TX = 64, SY_TY = 16;                // sample tile of pass 3
constexpr int SY_NY = SY_TY / 2 + 1;
constexpr int ZT = 16;                               // z-lines per workgroup in pass 2

constexpr int TPW = 4;  // tiles per workgroup (the next tile is prefetched into registers)

struct XYAnArgs {
  const void* in;           // storage ST::T, as every buffer of the wavelet kernels
  long long ip1, ip2, ivp;  // input pitches (y, z) and volume pitch; x pitch 1
  void* B;
  long long bx, by, bvp;    // band buffer row pitch (2 px), rows (2ncy) and volume pitch
  int px;                   // padded band width (ncx rounded up to 16): the detail band starts at x = px
  int ax, ay, az, ncx, ncy, tiles_x, groups_x;
  Db2F f;
};

constexpr int AN_NLD = (AN_W * AN_H + 255) / 256;

// grid.x = groups_x * tiles_y (a group is TPW consecutive x tiles of one tile row),
// grid.y = z * volumes
template <class ST>
__global__ void __launch_bounds__(256) k_wav_xy_an(XYAnArgs a) {
  using E = typename ST::T;
  __shared__ float2 s_in[AN_H * AN_W];
  __shared__ float2 s_x[AN_H * 2 * AN_TCX];
  const int tid = threadIdx.x;
  const int gx = blockIdx.x % a.groups_x;
  const int cy0 = (blockIdx.x / a.groups_x) * AN_TCY;
  const int z = blockIdx.y % a.az;
  const long long v = blockIdx.y / a.az;
  const E* src = (const E*)a.in + v * a.ivp + (long long)z * a.ip2;
  E* dst = (E*)a.B + v * a.bvp + (long long)z * a.bx * a.by;
  const int t0 = gx * TPW;
  const int nt = min(TPW, a.tiles_x - t0);
  E xn[AN_NLD];
  auto load = [&](int cx0) {
#pragma unroll
    for (int q = 0; q < AN_NLD; ++q) {
      int e = tid + 256 * q;
      if (e < AN_W * AN_H) {
        int i = e % AN_W, j = e / AN_W;
        xn[q] = src[(long long)sym(2 * cy0 - 2 + j, a.ay) * a.ip1 + sym(2 * cx0 - 2 + i, a.ax)];
      }
    }
  };
  load(t0 * AN_TCX);
  for (int it = 0; it < nt; ++it) {
    const int cx0 = (t0 + it) * AN_TCX;
#pragma unroll
    for (int q = 0; q < AN_NLD; ++q) {
      int e = tid + 256 * q;
      if (e < AN_W * AN_H) s_in[e] = ST::cvt(xn[q]);
    }
    __syncthreads();
    // along x: local sample 2c+3-k <-> global 2(cx0+c)+1-k
    for (int e = tid; e < AN_H * AN_TCX; e += 256) {
      int c = e % AN_TCX, j = e / AN_TCX;
      float2 sa = make_float2(0.f, 0.f), sd = make_float2(0.f, 0.f);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        float2 x = s_in[j * AN_W + 2 * c + 3 - k];
        fma2(sa, a.f.dlo[k], x);
        fma2(sd, a.f.dhi[k], x);
      }
      s_x[j * 2 * AN_TCX + c] = sa;
      s_x[j * 2 * AN_TCX + AN_TCX + c] = sd;
    }
    __syncthreads();
    if (it + 1 < nt) load(cx0 + AN_TCX);
    for (int e = tid; e < AN_TCY * 2 * AN_TCX; e += 256) {
      int c2 = e % (2 * AN_TCX), cy = e / (2 * AN_TCX);
      float2 sa = make_float2(0.f, 0.f), sd = make_float2(0.f, 0.f);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        float2 x = s_x[(2 * cy + 3 - k) * 2 * AN_TCX + c2];
        fma2(sa, a.f.dlo[k], x);
        fma2(sd, a.f.dhi[k], x);
      }
      int cxl = c2 < AN_TCX ? c2 : c2 - AN_TCX;
      if (cx0 + cxl < a.ncx && cy0 + cy < a.ncy) {
        long long gxo = (c2 < AN_TCX) ? cx0 + cxl : a.px + cx0 + cxl;
        ST::st(dst + (long long)(cy0 + cy) * a.bx + gxo, sa);
        ST::st(dst + (long long)(a.ncy + cy0 + cy) * a.bx + gxo, sd);
      }
    }
  }
}

struct ZArgs {
  void* B;
  long long bx, by, bvp;
  int az, ncz, ncx, ncy, px, tiles_x, groups_x;
  void* A;    // next-level input (ncx x ncy x ncz, row pitch pa) or null at the deepest level
  long long avp;
  int pa;
  int deepest;
  float alpha;
  Db2F f;
};

// grid.x = groups_x * by (TPW consecutive tiles of ZT x-neighbour z-lines at one Y), grid.y = volumes.
// NZR >= ceil(az / 16) values of a line are prefetched per thread.
template <int NZR, class ST>
__global__ void __launch_bounds__(256) k_wav_z(ZArgs a) {
  using E = typename ST::T;
  extern __shared__ float2 zs[];
  float2* L = zs;                    // az x ZT
  float2* Cf = zs + a.az * ZT;       // 2ncz x ZT
  const int tid = threadIdx.x;
  const int t = tid % ZT, r = tid / ZT;
  const int gx = blockIdx.x % a.groups_x;
  const int Y = blockIdx.x / a.groups_x;
  const long long v = blockIdx.y;
  const long long zp = a.bx * a.by;
  E* row = (E*)a.B + v * a.bvp + (long long)Y * a.bx;
  const int t0 = gx * TPW;
  const int nt = min(TPW, a.tiles_x - t0);
  E xn[NZR];
  auto load = [&](int X) {
    const bool act = X < a.bx && (X < a.px ? X : X - a.px) < a.ncx;
#pragma unroll
    for (int q = 0; q < NZR; ++q) {
      int z = r + (256 / ZT) * q;
      if (z < a.az) xn[q] = act ? row[X + z * zp] : ST::zero();
    }
  };
  load(t0 * ZT + t);
  for (int it = 0; it < nt; ++it) {
    const int X = (t0 + it) * ZT + t;
    const bool act = X < a.bx && (X < a.px ? X : X - a.px) < a.ncx;
#pragma unroll
    for (int q = 0; q < NZR; ++q) {
      int z = r + (256 / ZT) * q;
      if (z < a.az) L[z * ZT + t] = ST::cvt(xn[q]);
    }
    __syncthreads();
    const bool ll = X < a.ncx && Y < a.ncy;
    for (int c = r; c < a.ncz; c += 256 / ZT) {
      float2 sa = make_float2(0.f, 0.f), sd = make_float2(0.f, 0.f);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        float2 x = L[sym(2 * c + 1 - k, a.az) * ZT + t];
        fma2(sa, a.f.dlo[k], x);
        fma2(sd, a.f.dhi[k], x);
      }
      sd = shrink(sd, a.alpha);
      if (!ll || a.deepest) {
        sa = shrink(sa, a.alpha);
      } else {
        ST::st((E*)a.A + v * a.avp + (long long)c * a.pa * a.ncy + (long long)Y * a.pa + X, sa);
        sa = make_float2(0.f, 0.f);
      }
      Cf[c * ZT + t] = sa;
      Cf[(a.ncz + c) * ZT + t] = sd;
    }
    __syncthreads();
    if (it + 1 < nt) load(X + ZT);
    for (int s = r; s < a.az; s += 256 / ZT) {
      float2 acc = make_float2(0.f, 0.f);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        int idx = s + 2 - j;
        if (idx >= 0 && idx < 2 * a.ncz && (idx & 1) == 0) {
          int tt = idx >> 1;
          fma2(acc, a.f.rlo[j], Cf[tt * ZT + t]);
          fma2(acc, a.f.rhi[j], Cf[(a.ncz + tt) * ZT + t]);
        }
      }
      if (act) ST::st(row + X + s * zp, acc);
    }
    __syncthreads();  // L and Cf are rewritten by the next tile
  }
}

// Level-0 z pass of an iterate that is X_q = sum_k C_k conj(S[q][k]) at every
// voxel (the LLR step's output when its block grid covers the volume). The
// wavelet's analysis is linear with real filters, so the xy-bands of X_q are
// the same combination of the xy-bands of the K coefficient planes (pass 1 run
// on the planes, FP32). A workgroup keeps its ZT z-lines of all K plane bands
// in registers and, for every volume in turn, expands them (FP32), analyses
// the lines along z, shrinks and synthesises them back, writing the volume's
// bands (and the next level's input) exactly as k_wav_z would have: the
// back-projection and the per-volume xy analysis never touch HBM.
struct ZXArgs {
  const float2* BK;  // K planes of xy-bands, layout of B (FP32), plane pitch bkvp
  long long bkvp;
  const float2* S;   // nq x K, column-major
  int K, nq;
  ZArgs z;           // per-volume band buffer B (ST), next-level input A, geometry
};

constexpr int ZX_KM = 8;  // coefficient planes held per thread (K <= 8)

// grid.x = tiles_x * by: one tile of ZT x-neighbour z-lines at one Y, ZT * NCH
// threads. Thread (t = line, r = chunk) owns the CH consecutive samples z in
// [CH r, CH r + CH) of its line (CH even, NCH CH >= az, az a multiple of CH)
// and the coefficients c with 2c in that range (the last chunk also owns
// c = az/2): analysis needs the two samples before the chunk and synthesis the
// first coefficient after it, which come from the neighbouring chunks through
// LDS; everything else stays in registers.
template <int CH, int NCH, int KK, class ST>
__global__ void __launch_bounds__(ZT * NCH) k_wav_z_expand(ZXArgs g) {
  using E = typename ST::T;
  const ZArgs& a = g.z;
  constexpr int NC = CH / 2 + 1;                 // coefficient slots per thread (the last one: c = az/2)
  __shared__ float2 h_x[NCH][2][ZT];             // last two samples of each chunk
  __shared__ float2 h_c[NCH][2][ZT];             // first (lo, hi) coefficient of each chunk
  const int tid = threadIdx.x;
  const int t = tid % ZT, r = tid / ZT;
  const int Y = blockIdx.x / a.tiles_x;
  const int X = (blockIdx.x % a.tiles_x) * ZT + t;
  const bool act = X < a.bx && (X < a.px ? X : X - a.px) < a.ncx;
  const long long zp = a.bx * a.by;
  const int z0 = CH * r;
  const int nch = a.az / CH;                     // chunks holding samples (az is a multiple of CH)
  const bool live = r < nch;
  const bool last = r == nch - 1;
  // KK planes are held (KK = K, or 8 with planes K..7 zero)
  float2 kv[KK][CH];
#pragma unroll
  for (int k = 0; k < KK; ++k)
#pragma unroll
    for (int i = 0; i < CH; ++i)
      kv[k][i] = (k < g.K && live && act) ? g.BK[k * g.bkvp + (long long)Y * a.bx + X + (long long)(z0 + i) * zp]
                                          : make_float2(0.f, 0.f);
  const bool ll = X < a.ncx && Y < a.ncy;
  for (int v = 0; v < g.nq; ++v) {
    float2 sc[KK];
#pragma unroll
    for (int k = 0; k < KK; ++k) sc[k] = k < g.K ? g.S[v + (long long)k * g.nq] : make_float2(0.f, 0.f);
    // 1. this volume's samples of the chunk: x = sum_k C_k conj(S[v][k])
    float2 x[CH];
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      float2 acc = make_float2(0.f, 0.f);
#pragma unroll
      for (int k = 0; k < KK; ++k) acc = cadd(acc, cmulc(kv[k][i], sc[k]));
      x[i] = acc;
    }
    h_x[r][0][t] = x[CH - 2];
    h_x[r][1][t] = x[CH - 1];
    __syncthreads();  // halos of samples ready (and the previous volume's coefficient halos consumed)
    // samples z0 - 2 .. z0 + CH + 1 (symmetric extension at both ends)
    float2 w[CH + 4];
    if (r == 0) {
      w[0] = x[1];
      w[1] = x[0];
    } else {
      w[0] = h_x[r - 1][0][t];
      w[1] = h_x[r - 1][1][t];
    }
#pragma unroll
    for (int i = 0; i < CH; ++i) w[2 + i] = x[i];
    w[CH + 2] = x[CH - 1];  // used only by the last chunk: sym(az) = az - 1, sym(az + 1) = az - 2
    w[CH + 3] = x[CH - 2];
    // 2. analysis of c = z0/2 + m (m < CH/2), and c = az/2 for the last chunk:
    //    a[c] = sum_k dlo[k] x[sym(2c + 1 - k)], local sample index 2m + 3 - k
    float2 ca[NC], cd[NC];
    ca[NC - 1] = cd[NC - 1] = make_float2(0.f, 0.f);
    // slot NC - 1 (c = az/2) only for the last chunk; the others take it from the next chunk
#pragma unroll
    for (int m = 0; m < NC; ++m) {
      if (m == NC - 1 && !last) break;
      float2 sa = make_float2(0.f, 0.f), sd = make_float2(0.f, 0.f);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const float2 xv = w[2 * m + 3 - k];
        fma2(sa, a.f.dlo[k], xv);
        fma2(sd, a.f.dhi[k], xv);
      }
      const int c = z0 / 2 + m;
      sd = shrink(sd, a.alpha);
      if (!ll || a.deepest) {
        sa = shrink(sa, a.alpha);
      } else {
        if (live && (m < NC - 1 || last) && act)
          ST::st((E*)a.A + v * a.avp + (long long)c * a.pa * a.ncy + (long long)Y * a.pa + X, sa);
        sa = make_float2(0.f, 0.f);
      }
      ca[m] = sa;
      cd[m] = sd;
    }
    h_c[r][0][t] = ca[0];
    h_c[r][1][t] = cd[0];
    __syncthreads();  // coefficient halos ready; sample halos consumed
    // 3. synthesis of the chunk's samples: out[s] uses coefficients s/2 and s/2 + 1
    //    (the latter from the next chunk for the chunk's last two samples)
    if (live && !last) {  // r + 1 < nch <= NCH
      ca[CH / 2] = h_c[r + 1][0][t];
      cd[CH / 2] = h_c[r + 1][1][t];
    }
    if (live && act) {
      E* row = (E*)a.B + v * a.bvp + (long long)Y * a.bx + X;
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        // s = z0 + i; even s: taps j = 2 (c = s/2) and j = 0 (c = s/2 + 1); odd s: j = 3 and j = 1
        const int m = i / 2;
        float2 acc = make_float2(0.f, 0.f);
        if ((i & 1) == 0) {
          fma2(acc, a.f.rlo[2], ca[m]);
          fma2(acc, a.f.rhi[2], cd[m]);
          fma2(acc, a.f.rlo[0], ca[m + 1]);
          fma2(acc, a.f.rhi[0], cd[m + 1]);
        } else {
          fma2(acc, a.f.rlo[3], ca[m]);
          fma2(acc, a.f.rhi[3], cd[m]);
          fma2(acc, a.f.rlo[1], ca[m + 1]);
          fma2(acc, a.f.rhi[1], cd[m + 1]);
        }
        ST::st(row + (long long)(z0 + i) * zp, acc);
      }
    }
  }
}

struct XYSynArgs {
  const void* B;
  long long bx, by, bvp;
  int ncx, ncy, px;
  const void* Ap;    // reconstructed next level (ncx x ncy x ncz, row pitch pa) or null
  long long apvp;
  int pa;
  int ncz;           // this level's z coefficients per band (Ap's z length)
  void* out;
  long long op1, op2, ovp;
  int ax, ay, az, tiles_x, groups_x;
  int tpw;  // x tiles per workgroup
  Db2F f;
};

// Tiles are TX x SY_TY samples. Every index a thread touches is fixed by its thread id except for the tile's
// x origin, so the decompositions are done once per workgroup. Coefficients
// that do not exist (past ncx / ncy) are staged as zeros, so the synthesis
// reads the two taps of an output without range checks: output s uses
// coefficients floor(s/2) (tap 2 if s is even, 3 if odd) and floor(s/2) + 1
// (tap 0 / 1), tile origins being even.
template <class ST, int TX>
__global__ void __launch_bounds__(256) k_wav_xy_syn(XYSynArgs a) {
  using E = typename ST::T;
  constexpr int NX = TX / 2 + 1;                      // staged coefficient columns
  constexpr int NLD = (4 * SY_NY * NX + 255) / 256;   // band loads per thread
  constexpr int NYS = (2 * SY_TY * NX + 255) / 256;   // y-step outputs per thread
  constexpr int NXS = SY_TY * TX / 256;               // x-step outputs per thread
  constexpr int RP = 256 / TX;                        // output rows per x-step pass
  constexpr int NAQ = (SY_NY * NX + 255) / 256;       // load slots that can hold the LL band (b = 0)
  __shared__ float2 band[4 * SY_NY * NX];   // [xb + 2*yb][iy][ix]
  __shared__ float2 yrow[2 * SY_TY * NX];   // [xb][sy][ix]
  const int tid = threadIdx.x;
  const int gx = blockIdx.x % a.groups_x;
  const int y0 = (blockIdx.x / a.groups_x) * SY_TY;
  const int z = blockIdx.y % a.az;
  const long long v = blockIdx.y / a.az;
  const int ty0 = y0 / 2;
  const E* src = (const E*)a.B + v * a.bvp + (long long)z * a.bx * a.by;
  const E* ap = a.Ap ? (const E*)a.Ap + v * a.apvp : nullptr;
  // the LL band of this z slice picks up syn_z(Ap, 0): taps j with z+2-j even and in range
  int za = -1, zb = -1;
  float wa = 0.f, wb = 0.f;
  if (ap) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      int idx = z + 2 - j;
      if (idx >= 0 && idx < 2 * a.ncz && (idx & 1) == 0) {
        if (za < 0) {
          za = idx >> 1;
          wa = a.f.rlo[j];
        } else {
          zb = idx >> 1;
          wb = a.f.rlo[j];
        }
      }
    }
  }
  const int pza = za * a.pa * a.ncy, pzb = zb * a.pa * a.ncy;
  // load slots: band element e = tid + 256 q -> row offset in the slice (or -1) and column
  int lofs[NLD], lix[NLD], aofs[NLD];
#pragma unroll
  for (int q = 0; q < NLD; ++q) {
    const int e = tid + 256 * q;
    lofs[q] = -1;
    aofs[q] = -1;
    lix[q] = 0;
    if (e < 4 * SY_NY * NX) {
      const int ix = e % NX, rem = e / NX;
      const int iy = rem % SY_NY, b = rem / SY_NY;
      const int cy = ty0 + iy;
      lix[q] = ix;
      if (cy < a.ncy) {
        lofs[q] = ((b >> 1) ? a.ncy + cy : cy) * (int)a.bx + ((b & 1) ? a.px : 0);
        if (b == 0 && ap) aofs[q] = cy * a.pa;
      }
    }
  }
  // y step: output e -> (xb, sy, ix); its two taps read band rows t_a, t_a + 1
  // with the weights of the output's parity (bit q of yodd)
  int yo[NYS];
  unsigned yodd = 0;
#pragma unroll
  for (int q = 0; q < NYS; ++q) {
    const int e = min(tid + 256 * q, 2 * SY_TY * NX - 1);
    const int ix = e % NX, rem = e / NX;
    const int sy = rem % SY_TY, xb = rem / SY_TY;
    yo[q] = (xb * SY_NY + (sy >> 1)) * NX + ix;
    yodd |= (unsigned)(sy & 1) << q;
  }
  constexpr int HB = 2 * SY_NY * NX;  // lo band rows -> hi band rows
  // x step: output e = tid + 256 q -> (sy = tid/TX + RP q, sx = tid % TX)
  const int sx = tid % TX, sy0 = tid / TX, xodd = sx & 1;
  const float xl0 = a.f.rlo[xodd ? 3 : 2], xl1 = a.f.rlo[xodd ? 1 : 0];
  const float xh0 = a.f.rhi[xodd ? 3 : 2], xh1 = a.f.rhi[xodd ? 1 : 0];
  const int t0 = gx * a.tpw;
  const int nt = min(a.tpw, a.tiles_x - t0);
  // raw words of the next tile: the band value and, for LL slots (q < NAQ),
  // the two taps of the next level (zero where absent: weight x 0 adds 0)
  E bn[NLD], ba[NAQ], bb[NAQ];
  auto load = [&](int x0) {
    const int tx0 = x0 / 2;
#pragma unroll
    for (int q = 0; q < NLD; ++q) {
      const int cx = tx0 + lix[q];
      bn[q] = ST::zero();
      if (q < NAQ) {
        ba[q] = ST::zero();
        bb[q] = ST::zero();
      }
      if (lofs[q] >= 0 && cx < a.ncx) {
        bn[q] = src[lofs[q] + cx];  // (cached: neighbouring tiles share band rows and columns)
        if (q < NAQ && aofs[q] >= 0) {
          if (za >= 0) ba[q] = ap[pza + aofs[q] + cx];
          if (zb >= 0) bb[q] = ap[pzb + aofs[q] + cx];
        }
      }
    }
  };
  load(t0 * TX);
  E* dst = (E*)a.out + v * a.ovp + (long long)z * a.op2;
  for (int it = 0; it < nt; ++it) {
    const int x0 = (t0 + it) * TX;
#pragma unroll
    for (int q = 0; q < NLD; ++q) {
      const int e = tid + 256 * q;
      if (e < 4 * SY_NY * NX) {
        float2 b = ST::cvt(bn[q]);
        if (q < NAQ) {
          fma2(b, wa, ST::cvt(ba[q]));
          fma2(b, wb, ST::cvt(bb[q]));
        }
        band[e] = b;
      }
    }
    __syncthreads();
    if (it + 1 < nt) load(x0 + TX);
#pragma unroll
    for (int q = 0; q < NYS; ++q) {
      const int e = tid + 256 * q;
      if (e < 2 * SY_TY * NX) {
        const bool odd = (yodd >> q) & 1;
        const float l0 = odd ? a.f.rlo[3] : a.f.rlo[2], l1 = odd ? a.f.rlo[1] : a.f.rlo[0];
        const float h0 = odd ? a.f.rhi[3] : a.f.rhi[2], h1 = odd ? a.f.rhi[1] : a.f.rhi[0];
        const int o = yo[q];
        float2 acc = make_float2(0.f, 0.f);
        fma2(acc, l0, band[o]);
        fma2(acc, h0, band[o + HB]);
        fma2(acc, l1, band[o + NX]);
        fma2(acc, h1, band[o + HB + NX]);
        yrow[e] = acc;
      }
    }
    __syncthreads();
    const int gxo = x0 + sx;
#pragma unroll
    for (int q = 0; q < NXS; ++q) {
      const int sy = sy0 + RP * q, gyo = y0 + sy;
      if (gxo >= a.ax || gyo >= a.ay) continue;
      const int o = sy * NX + (sx >> 1);
      float2 acc = make_float2(0.f, 0.f);
      fma2(acc, xl0, yrow[o]);
      fma2(acc, xh0, yrow[o + SY_TY * NX]);
      fma2(acc, xl1, yrow[o + 1]);
      fma2(acc, xh1, yrow[o + 1 + SY_TY * NX]);
      ST::st(dst + (long long)gyo * a.op1 + gxo, acc);
    }
    __syncthreads();  // band and yrow are rewritten by the next tile
  }
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


static inline size_t pad16(size_t n) { return (n + 15) / 16 * 16; }

// Per-volume workspace, with rows padded to 16 elements (128 B) so every band
// row segment is cache-line aligned: the band buffer of every level
// (2*pad16(ncx) x 2ncy x nz) and the next-level inputs of levels 1..L-1
// (pad16(ncx) x ncy x ncz of the level above).
size_t wavelet_workspace_elems(const size_t* dims, int L, size_t* coef_elems, size_t* tmp_elems) {
  std::vector<std::array<size_t, 3>> in_dims, cdims;
  wavelet_geometry(dims, L, in_dims, cdims);
  size_t c = 0;
  for (int l = 0; l < L; ++l) {
    c += 2 * pad16(cdims[l][0] / 2) * cdims[l][1] * in_dims[l][2];
    if (l > 0) c += pad16(in_dims[l][0]) * in_dims[l][1] * in_dims[l][2];
  }
  *coef_elems = c;
  *tmp_elems = 0;
  return c;
}

template <class ST>
static llr_status wavelet_run(llr_ctx* ctx, void* vols_v, size_t nvols, const size_t* dims, float alpha, int L,
                              void* coef_v, const WaveletFromSubspace* sub) {
  using E = typename ST::T;
  E* vols = (E*)vols_v;
  Db2F f = db2_filters();
  size_t N = dims[0] * dims[1] * dims[2];
  if (L <= 0) {  // the volume is its own approximation
    hipLaunchKernelGGL(k_shrink_store<ST>, dim3(grid_for(N * nvols, 256)), dim3(256), 0, ctx->stream, vols,
                       (long long)(N * nvols), alpha);
    LLR_HIP(hipGetLastError());
    return LLR_OK;
  }
  std::vector<std::array<size_t, 3>> in_dims, cdims;
  wavelet_geometry(dims, L, in_dims, cdims);
  std::vector<E*> Bp(L), Ap(L, nullptr);
  std::vector<size_t> bvol(L), avol(L, 0), px(L), pa(L, 0);
  {
    E* p = (E*)coef_v;
    for (int l = 0; l < L; ++l) {
      px[l] = pad16(cdims[l][0] / 2);
      bvol[l] = 2 * px[l] * cdims[l][1] * in_dims[l][2];
      Bp[l] = p;
      p += bvol[l] * nvols;
    }
    for (int l = 1; l < L; ++l) {
      pa[l] = pad16(in_dims[l][0]);
      avol[l] = pa[l] * in_dims[l][1] * in_dims[l][2];
      Ap[l] = p;
      p += avol[l] * nvols;
    }
  }
  hipStream_t s = ctx->stream;
  for (int l = 0; l < L; ++l) {
    const size_t* in = in_dims[l].data();
    const size_t* cd = cdims[l].data();
    int ncx = (int)cd[0] / 2, ncy = (int)cd[1] / 2, ncz = (int)cd[2] / 2;
    const long long bx = 2 * (long long)px[l], by = (long long)cd[1];
    const bool fused = l == 0 && sub != nullptr;
    // pass 1 (on the K coefficient planes, FP32, when fused)
    XYAnArgs a1{};
    a1.in = fused ? (const void*)sub->C : l == 0 ? (const void*)vols : (const void*)Ap[l];
    a1.ip1 = l == 0 ? (long long)in[0] : (long long)pa[l];
    a1.ip2 = a1.ip1 * (long long)in[1];
    a1.ivp = a1.ip2 * (long long)in[2];
    a1.B = fused ? (void*)sub->BK : (void*)Bp[l];
    a1.bx = bx;
    a1.by = by;
    a1.bvp = (long long)bvol[l];
    a1.px = (int)px[l];
    a1.ax = (int)in[0];
    a1.ay = (int)in[1];
    a1.az = (int)in[2];
    a1.ncx = ncx;
    a1.ncy = ncy;
    a1.tiles_x = (ncx + AN_TCX - 1) / AN_TCX;
    a1.groups_x = (a1.tiles_x + TPW - 1) / TPW;
    a1.f = f;
    int tiles_y = (ncy + AN_TCY - 1) / AN_TCY;
    dim3 g1((unsigned)(a1.groups_x * tiles_y), (unsigned)(in[2] * (fused ? (size_t)sub->K : nvols)));
    {
      KernelTimer kt(ctx, l == 0 ? kStageWavIn : kStageWavDeep);
      if (fused)
        hipLaunchKernelGGL(k_wav_xy_an<StoreF32>, g1, dim3(256), 0, s, a1);
      else
        hipLaunchKernelGGL(k_wav_xy_an<ST>, g1, dim3(256), 0, s, a1);
    }
    LLR_HIP(hipGetLastError());
    // pass 2
    ZArgs a2{};
    a2.B = Bp[l];
    a2.bx = bx;
    a2.by = by;
    a2.bvp = (long long)bvol[l];
    a2.az = (int)in[2];
    a2.ncz = ncz;
    a2.ncx = ncx;
    a2.ncy = ncy;
    a2.px = (int)px[l];
    a2.tiles_x = (int)((bx + ZT - 1) / ZT);
    a2.groups_x = (a2.tiles_x + TPW - 1) / TPW;
    a2.deepest = l == L - 1;
    a2.A = a2.deepest ? nullptr : Ap[l + 1];
    a2.avp = a2.deepest ? 0 : (long long)avol[l + 1];
    a2.pa = a2.deepest ? 0 : (int)pa[l + 1];
    a2.alpha = alpha;
    a2.f = f;
    size_t smem = (size_t)(in[2] + 2 * ncz) * ZT * sizeof(float2);  // LDS holds FP32 values
    if (smem > 160 * 1024 || in[2] > 512) return fail(LLR_ERR_UNSUPPORTED, "z length too large for the wavelet z pass");
    int nzr = (int)((in[2] + (256 / ZT) - 1) / (256 / ZT));
    if (fused) {
      if (sub->K > ZX_KM) return fail(LLR_ERR_UNSUPPORTED, "fused wavelet: rank <= 8");
      ZXArgs ax{};
      ax.BK = sub->BK;
      ax.bkvp = (long long)bvol[0];
      ax.S = sub->S;
      ax.K = sub->K;
      ax.nq = (int)nvols;
      ax.z = a2;
      dim3 gx((unsigned)(a2.tiles_x * by));
      // chunks of 2 (az <= 32 with 16 chunks) or 4 samples (16 or 32 chunks)
      const int nch = in[2] <= 64 ? 16 : 32;
      const int ch = 2 * (int)((in[2] + 2 * nch - 1) / (2 * nch));
      if (in[2] % 2 || in[2] % ch || ch > 4) return fail(LLR_ERR_UNSUPPORTED, "fused wavelet: z length");
      KernelTimer kt(ctx, kStageWavMid);
#define ZX_LAUNCH(C, N)                                                                   \
  if (sub->K == 7)                                                                        \
    hipLaunchKernelGGL((k_wav_z_expand<C, N, 7, ST>), gx, dim3(ZT * N), 0, s, ax);       \
  else                                                                                    \
    hipLaunchKernelGGL((k_wav_z_expand<C, N, ZX_KM, ST>), gx, dim3(ZT * N), 0, s, ax);
      if (nch == 16 && ch == 2) {
        ZX_LAUNCH(2, 16)
      } else if (nch == 16) {
        ZX_LAUNCH(4, 16)
      } else if (ch == 2) {
        ZX_LAUNCH(2, 32)
      } else {
        ZX_LAUNCH(4, 32)
      }
#undef ZX_LAUNCH
      LLR_HIP(hipGetLastError());
      continue;
    }
    dim3 g2((unsigned)(a2.groups_x * by), (unsigned)nvols);
    KernelTimer kt(ctx, l == 0 ? kStageWavMid : kStageWavDeep);
#define WZ_CASE(NN)                                                                                       \
  if (nzr <= NN) {                                                                                        \
    if (smem > 64 * 1024)                                                                                 \
      LLR_HIP(hipFuncSetAttribute((const void*)k_wav_z<NN, ST>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                                  (int)smem));                                                            \
    hipLaunchKernelGGL((k_wav_z<NN, ST>), g2, dim3(256), smem, s, a2);                                    \
  } else
    WZ_CASE(1) WZ_CASE(2) WZ_CASE(4) WZ_CASE(8) WZ_CASE(16) WZ_CASE(32) {}
#undef WZ_CASE
    LLR_HIP(hipGetLastError());
  }
  for (int l = L - 1; l >= 0; --l) {
    const size_t* od = in_dims[l].data();
    const size_t* cd = cdims[l].data();
    XYSynArgs a3{};
    a3.B = Bp[l];
    a3.bx = 2 * (long long)px[l];
    a3.by = (long long)cd[1];
    a3.bvp = (long long)bvol[l];
    a3.ncx = (int)cd[0] / 2;
    a3.ncy = (int)cd[1] / 2;
    a3.px = (int)px[l];
    a3.ncz = (int)cd[2] / 2;
    a3.Ap = (l < L - 1) ? Ap[l + 1] : nullptr;
    a3.apvp = (l < L - 1) ? (long long)avol[l + 1] : 0;
    a3.pa = (l < L - 1) ? (int)pa[l + 1] : 0;
    a3.out = l == 0 ? vols : Ap[l];
    a3.op1 = l == 0 ? (long long)od[0] : (long long)pa[l];
    a3.op2 = a3.op1 * (long long)od[1];
    a3.ovp = a3.op2 * (long long)od[2];
    a3.ax = (int)od[0];
    a3.ay = (int)od[1];
    a3.az = (int)od[2];
    constexpr int TXS = SY_TX;  // (128 for half storage measured slower)
    a3.tiles_x = (int)((od[0] + TXS - 1) / TXS);
    a3.tpw = TPW;
    a3.groups_x = (a3.tiles_x + a3.tpw - 1) / a3.tpw;
    a3.f = f;
    int tiles_y = (int)((od[1] + SY_TY - 1) / SY_TY);
    dim3 g3((unsigned)(a3.groups_x * tiles_y), (unsigned)(od[2] * nvols));
    {
      KernelTimer kt(ctx, l == 0 ? kStageWavOut : kStageWavDeep);
      hipLaunchKernelGGL((k_wav_xy_syn<ST, TXS>), g3, dim3(256), 0, s, a3);
    }
    LLR_HIP(hipGetLastError());
  }
  return LLR_OK;
}

// In-place wavelet soft threshold of nvols volumes; coef holds nvols *
// coef_elems elements of the same storage (t1, t2 unused).
llr_status wavelet_denoise_device(llr_ctx* ctx, void* vols, size_t nvols, const size_t* dims, float alpha, int L,
                                  void* coef, void* t1, void* t2, bool h16) {
  (void)t1;
  (void)t2;
  return h16 ? wavelet_run<StoreF16>(ctx, vols, nvols, dims, alpha, L, coef, nullptr)
              : wavelet_run<StoreF32>(ctx, vols, nvols, dims, alpha, L, coef, nullptr);
}

size_t wavelet_subspace_elems(const size_t* dims, int L, int K) {
  if (L <= 0) return 0;
  std::vector<std::array<size_t, 3>> in_dims, cdims;
  wavelet_geometry(dims, L, in_dims, cdims);
  return 2 * pad16(cdims[0][0] / 2) * cdims[0][1] * in_dims[0][2] * (size_t)K;
}

bool wavelet_subspace_supported(const size_t* dims, int L, int K) {
  const size_t nz = dims[2], nch = nz <= 64 ? 16 : 32, ch = 2 * ((nz + 2 * nch - 1) / (2 * nch));
  return L >= 1 && K >= 1 && K <= ZX_KM && nz >= 4 && nz <= 128 && nz % 2 == 0 && nz % ch == 0;
}

llr_status wavelet_denoise_subspace(llr_ctx* ctx, void* vols, size_t nvols, const size_t* dims, float alpha, int L,
                                    void* coef, const WaveletFromSubspace& sub, bool h16) {
  if (!wavelet_subspace_supported(dims, L, sub.K)) return fail(LLR_ERR_UNSUPPORTED, "fused subspace wavelet");
  return h16 ? wavelet_run<StoreF16>(ctx, vols, nvols, dims, alpha, L, coef, &sub)
             : wavelet_run<StoreF32>(ctx, vols, nvols, dims, alpha, L, coef, &sub);
}

}  // namespace llr
