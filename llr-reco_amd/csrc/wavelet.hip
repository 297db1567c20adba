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

// Wavelet of an iterate that is X_q = sum_k C_k conj(S[q][k]) at every voxel
// (the LLR step's output when its block grid covers the volume). Analysis is
// linear with real filters, so every level's coefficients of X_q are the same
// combination of the coefficients of the K planes C_k: every analysis pass
// runs on the K planes only (FP32), and each volume is only synthesised. The
// per-volume z synthesis expands its coefficients from the K planes'
// (cmulc with the volume's row of S), shrinks them and synthesises the z-lines
// of the xy-bands; the xy synthesis then runs as in the plain path. Levels are
// synthesised coarse to fine, so a level's LL band reads the reconstructed
// next level as its z-approximation (not shrunk) and the xy synthesis has no
// next-level taps to add. The back-projection and every per-volume analysis
// never touch HBM.

// K-plane analysis of one level in one pass (the xy-bands of a slice are never
// stored): a workgroup owns one AN_TCX x AN_TCY coefficient tile of one plane
// and AN3_NS consecutive z slots c. It walks the logical z slices
// i = 2 c0 - 2 .. 2 c_end + 1, each loaded through the symmetric extension
// (slice sym(i)), filters it along x and y as k_wav_xy_an does and keeps the
// last four slices' xy-band values in registers: after an odd slice i = 2c + 1
// they are the slices 2c + 1 - j of the z filter's taps j. Slot c's z
// coefficients go to WK, a box bx x by x 2ncz per plane (xy-band position
// (X, Y), approximation at slot c, detail at ncz + c), except the LL band's
// approximation above the deepest level, which is the next level's input (AK,
// ncx x ncy x ncz, row pitch pa).
// At 256x256x128 (K = 7) this replaces an xy pass into a band box and a z pass
// out of it (0.28 + 0.21 ms, 1.07 GB of band traffic).
constexpr int AN3_NS = 8;
constexpr int AN3_TH = 512;  // threads: two y-step outputs (four window values per slice) each
constexpr int AN3_NLD = (AN_W * AN_H + AN3_TH - 1) / AN3_TH;

struct An3Args {
  const float2* in;         // K planes (ax x ay x az, row pitch ip1, plane pitch ivp)
  long long ip1, ip2, ivp;
  float2* WK;               // box bx x by x 2ncz per plane
  long long wvp;
  float2* AK;               // next level's input (ncx x ncy x ncz, row pitch pa), per plane
  long long akvp;
  int pa;
  long long bx, by;
  int px, ax, ay, az, ncx, ncy, ncz, tiles_x, deepest;
  Db2F f;
};

// grid = (tiles_x * tiles_y, ceil(ncz / AN3_NS), K)
__global__ void __launch_bounds__(AN3_TH) k_wav_an3_k(An3Args a) {
  __shared__ float2 s_in[AN_H * AN_W];
  __shared__ float2 s_x[AN_H * 2 * AN_TCX];
  const int tid = threadIdx.x;
  const int cx0 = (int)(blockIdx.x % a.tiles_x) * AN_TCX;
  const int cy0 = (int)(blockIdx.x / a.tiles_x) * AN_TCY;
  const int c0 = (int)blockIdx.y * AN3_NS;
  const int c1 = min(c0 + AN3_NS, a.ncz);
  const int k = blockIdx.z;
  const float2* src = a.in + k * a.ivp;
  // this thread's y-step outputs (e = tid + AN3_TH q): band column c2, local row cy
  constexpr int NQ = AN_TCY * 2 * AN_TCX / AN3_TH;
  float2 win[4][NQ][2];  // [slice 2c - 2 .. 2c + 1][q][y lo, y hi]
  float2 xn[AN3_NLD];
  auto load = [&](int i) {
    const float2* p = src + (long long)sym(i, a.az) * a.ip2;
#pragma unroll
    for (int q = 0; q < AN3_NLD; ++q) {
      const int e = tid + AN3_TH * q;
      if (e < AN_W * AN_H) {
        const int ii = e % AN_W, jj = e / AN_W;
        xn[q] = p[(long long)sym(2 * cy0 - 2 + jj, a.ay) * a.ip1 + sym(2 * cx0 - 2 + ii, a.ax)];
      }
    }
  };
  const int i0 = 2 * c0 - 2, i1 = 2 * c1 - 1;
  load(i0);
  for (int i = i0; i <= i1; ++i) {
#pragma unroll
    for (int q = 0; q < AN3_NLD; ++q) {
      const int e = tid + AN3_TH * q;
      if (e < AN_W * AN_H) s_in[e] = xn[q];
    }
    __syncthreads();
    for (int e = tid; e < AN_H * AN_TCX; e += AN3_TH) {
      const int c = e % AN_TCX, j = e / AN_TCX;
      float2 sa = make_float2(0.f, 0.f), sd = make_float2(0.f, 0.f);
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const float2 x = s_in[j * AN_W + 2 * c + 3 - t];
        fma2(sa, a.f.dlo[t], x);
        fma2(sd, a.f.dhi[t], x);
      }
      s_x[j * 2 * AN_TCX + c] = sa;
      s_x[j * 2 * AN_TCX + AN_TCX + c] = sd;
    }
    __syncthreads();
    if (i < i1) load(i + 1);
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        win[s][q][0] = win[s + 1][q][0];
        win[s][q][1] = win[s + 1][q][1];
      }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int e = tid + AN3_TH * q, c2 = e % (2 * AN_TCX), cy = e / (2 * AN_TCX);
      float2 sa = make_float2(0.f, 0.f), sd = make_float2(0.f, 0.f);
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const float2 x = s_x[(2 * cy + 3 - t) * 2 * AN_TCX + c2];
        fma2(sa, a.f.dlo[t], x);
        fma2(sd, a.f.dhi[t], x);
      }
      win[3][q][0] = sa;
      win[3][q][1] = sd;
    }
    if (((i - i0) & 1) == 0 || i < 2 * c0 + 1) continue;  // slot c = (i - 1) / 2 once its four slices are in
    const int c = (i - 1) / 2;
    const long long zp = a.bx * a.by;
    float2* wk = a.WK + k * a.wvp;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int e = tid + AN3_TH * q, c2 = e % (2 * AN_TCX), cy = e / (2 * AN_TCX);
      const int cxl = c2 < AN_TCX ? c2 : c2 - AN_TCX;
      if (cx0 + cxl >= a.ncx || cy0 + cy >= a.ncy) continue;
      const int X = c2 < AN_TCX ? cx0 + cxl : a.px + cx0 + cxl;
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const int Y = b == 0 ? cy0 + cy : a.ncy + cy0 + cy;
        float2 zl = make_float2(0.f, 0.f), zh = make_float2(0.f, 0.f);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          fma2(zl, a.f.dlo[t], win[3 - t][q][b]);
          fma2(zh, a.f.dhi[t], win[3 - t][q][b]);
        }
        float2* w = wk + (long long)Y * a.bx + X;
        if (X < a.ncx && Y < a.ncy && !a.deepest)
          a.AK[k * a.akvp + ((long long)c * a.ncy + Y) * a.pa + X] = zl;
        else
          w[(long long)c * zp] = zl;
        w[(long long)(a.ncz + c) * zp] = zh;
      }
    }
  }
}

// Per-volume z synthesis of one level from the K planes' z coefficients.
struct ZSynArgs {
  const float2* WK;  // K planes of z coefficients (FP32), box bx x by x 2ncz, plane pitch wvp
  long long wvp;
  const float2* S;   // nq x K, column-major
  int K, nq;
  const void* Ar;    // reconstructed next level per volume (ST; ncx x ncy x ncz, row pitch pa), null at the deepest level
  long long arvp;
  int pa;
  void* B;           // per-volume xy-bands (ST), box bx x by x az, volume pitch bvp; rows ordered (Y, z):
  long long bvp;     // element (X, Y, z) at (Y az + z) bx + X
  long long bx, by;
  int az, ncz, ncx, ncy, px, deepest;
  float alpha;
  Db2F f;
};

// out[s] from the slots floor(s/2) (lo0, hi0) and floor(s/2) + 1 (lo1, hi1):
// taps 2 and 0 for even s, 3 and 1 for odd s (kernels.cu:230-261)
__device__ __forceinline__ float2 zsyn_tap(const Db2F& f, bool odd, float2 lo0, float2 hi0, float2 lo1, float2 hi1) {
  float2 acc = make_float2(0.f, 0.f);
  fma2(acc, f.rlo[odd ? 3 : 2], lo0);
  fma2(acc, f.rhi[odd ? 3 : 2], hi0);
  fma2(acc, f.rlo[odd ? 1 : 0], lo1);
  fma2(acc, f.rhi[odd ? 1 : 0], hi1);
  return acc;
}

constexpr int ZS_KM = 8;  // coefficient planes (K <= 8)

// Per volume q, slot c of a band position holds (lo, hi) = shrink(sum_k
// W_k conj(S[q][k])) of the K planes' z coefficients W_k, except the LL band's
// z-approximation above the deepest level, which is the reconstructed next
// level as it is; z sample s is then the synthesis of slots floor(s/2) and
// floor(s/2) + 1. The expansion runs on the matrix cores: for one band slot
// of 16 columns, the expansion into 16 volumes is a real matrix product:
// E[q][X] = sum_k S[q][k] W_k[X] for each of Re W and Im W when S is real; a
// complex S takes four more k-steps, E_re = sum_k (W.x S.x + W.y S.y), E_im =
// sum_k (W.y S.x - W.x S.y). v_mfma_f32_16x16x4_f32 takes A (16 x 4) with
// lane l holding A[l % 16][l / 16], B (4 x 16) with lane l holding
// B[l / 16][l % 16], and leaves D[4 (l / 16) + i][l % 16] in lane l, i < 4.
// A wave owns 16 columns (lane column X0 + l % 16) of one band row Y and the
// samples [2 c0, 2 c0 + 2 CS) of z: slots c0 .. c0 + CS. It holds those slots'
// K-plane values as B operands (planes l / 16 and l / 16 + 4, padded to 8
// with zeros) and walks the volumes 16 at a time (A: the S rows of 16
// volumes). Each lane then has its column's (lo, hi) pair of every slot for
// volumes q0 + 4 (l / 16) + i, shrinks them and synthesises its z samples slot
// after slot: no value leaves the lane. Workgroup: 4 waves, 2 column tiles x 2
// z chunks; grid = (column-tile pairs x by x chunk pairs, volume ranges), the
// column-tile pairs fastest when TF is set. NT: non-temporal stores.
constexpr int ZM_CS = 8;  // slots per z chunk (2 ZM_CS samples)

typedef float zm_v4f __attribute__((ext_vector_type(4)));

template <int CS, bool RS, class ST, bool NT, bool TF>
__global__ void __launch_bounds__(256, 3) k_wav_zsyn_m(ZSynArgs a, const float2* __restrict__ S, int nchunk, int qper) {
  using E = typename ST::T;
  constexpr int NS = RS ? 2 : 4;  // k-steps of 4 planes
  // the wave index is uniform: as a scalar, every slot and sample offset is too
  const int lane = (int)threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int jl = lane & 15, kq = lane >> 4;
  const int nchunk2 = (nchunk + 1) >> 1;
  int b = (int)blockIdx.x;
  int cp, Y, tp;
  if (TF) {  // column-tile pairs fastest
    const int ntp = (int)((a.bx + 31) / 32);
    tp = b % ntp;
    b /= ntp;
    cp = b % nchunk2;
    Y = b / nchunk2;
  } else {
    cp = b % nchunk2;
    b /= nchunk2;
    Y = b % (int)a.by;
    tp = b / (int)a.by;
  }
  const int X0 = (2 * tp + (wv & 1)) * 16;
  const int ch = 2 * cp + (wv >> 1);
  if (X0 >= a.bx || ch >= nchunk) return;  // the whole wave (bx is a multiple of 16)
  const int X = X0 + jl;
  const bool act = (X < a.px ? X : X - a.px) < a.ncx;
  // the LL band's z-approximation is the reconstructed next level; px is a
  // multiple of 16, so this holds for every active lane of the tile or none
  const bool lo_ar = X0 < a.ncx && Y < a.ncy && !a.deepest;
  const int c0 = ch * CS;
  const long long zp = a.bx * a.by;
  const float2* wk = a.WK + (long long)Y * a.bx + X;
  float2 wl[CS + 1][2], wh[CS + 1][2];
#pragma unroll
  for (int t = 0; t <= CS; ++t)
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int kk = kq + 4 * s, c = c0 + t;
      const bool on = act && kk < a.K && c < a.ncz;
      wl[t][s] = (on && !lo_ar) ? wk[kk * a.wvp + (long long)c * zp] : make_float2(0.f, 0.f);
      wh[t][s] = on ? wk[kk * a.wvp + (long long)(a.ncz + c) * zp] : make_float2(0.f, 0.f);
    }
  const int qbeg = (int)blockIdx.y * qper, qend = min(a.nq, qbeg + qper);
  const long long arz = (long long)a.pa * a.ncy;
  // the next level's rows are pa >= px > X long, so an inactive lane's read
  // stays inside its row
  const E* arb = lo_ar ? (const E*)a.Ar + (long long)Y * a.pa + X : nullptr;
  E* ob = (E*)a.B + (long long)Y * a.az * a.bx + X;
  // Loads are issued a step ahead of their use: the memory counter a wave waits
  // on is shared by loads and stores and completes in order, so a load issued
  // after a slot's stores would also wait for those stores to be written.
  // S rows of a volume group (the A operand), one group ahead:
  auto load_s = [&](int q0, float2(&sv)[2]) {
    const int qa = q0 + jl;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int kk = kq + 4 * s;
      sv[s] = (qa < qend && kk < a.K) ? S[qa + (long long)kk * a.nq] : make_float2(0.f, 0.f);
    }
  };
  // next-level values of slot c for this lane's volumes q0 + 4 kq + i, one slot
  // ahead (volumes past qend, only in the last group, read clamped and not stored)
  auto load_ar = [&](int q0, int c, E(&an)[4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int q = min(q0 + 4 * kq + i, qend - 1);
      an[i] = c < a.ncz ? arb[(long long)q * a.arvp + c * arz] : ST::zero();
    }
  };
  float2 svn[2];
  load_s(qbeg, svn);
  E ann[4];
  if (lo_ar) load_ar(qbeg, c0, ann);
  for (int q0 = qbeg; q0 < qend; q0 += 16) {
    float sa[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) sa[s] = s < 2 ? svn[s].x : svn[s - 2].y;
    const int qn = q0 + 16;
    if (qn < qend) load_s(qn, svn);
    const int qi = q0 + 4 * kq;
    bool qok[4];
    E* oq = ob + (long long)qi * a.bvp;
    // per group, not hoisted: the slot offsets below are then scalars added here
    asm volatile("" : "+v"(oq));
#pragma unroll
    for (int i = 0; i < 4; ++i) qok[i] = act && qi + i < qend;
    E an[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) an[i] = ann[i];
    float2 plo[4], phi[4];
#pragma unroll
    for (int t = 0; t <= CS; ++t) {
      const int c = c0 + t;
      // keeps the slots in order (the compiler would otherwise start every
      // slot's products at once and hold all their results)
#pragma unroll
      for (int s = 0; s < NS; ++s) asm volatile("" : "+v"(sa[s])::"memory");
      E anx[4];
      if (lo_ar) {
        if (t < CS)
          load_ar(q0, c + 1, anx);
        else if (qn < qend)
          load_ar(qn, c0, anx);
      }
      zm_v4f hr = {0.f, 0.f, 0.f, 0.f}, hm = hr, lr = hr, lm = hr;
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        const float2 w = wh[t][s & 1];
        hr = __builtin_amdgcn_mfma_f32_16x16x4f32(sa[s], s < 2 ? w.x : w.y, hr, 0, 0, 0);
        hm = __builtin_amdgcn_mfma_f32_16x16x4f32(sa[s], s < 2 ? w.y : -w.x, hm, 0, 0, 0);
      }
      if (!lo_ar) {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          const float2 w = wl[t][s & 1];
          lr = __builtin_amdgcn_mfma_f32_16x16x4f32(sa[s], s < 2 ? w.x : w.y, lr, 0, 0, 0);
          lm = __builtin_amdgcn_mfma_f32_16x16x4f32(sa[s], s < 2 ? w.y : -w.x, lm, 0, 0, 0);
        }
      }
      float2 lo[4], hi[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        hi[i] = shrink(make_float2(hr[i], hm[i]), a.alpha);
        lo[i] = lo_ar ? ST::cvt(an[i]) : shrink(make_float2(lr[i], lm[i]), a.alpha);
      }
      if (t > 0) {
        const int z = 2 * (c - 1);  // samples z, z + 1 from slots c - 1, c
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          if (!qok[i]) continue;
          E* o = oq + (i * a.bvp + (long long)z * a.bx);
          if (NT) {
            if (z < a.az) ST::stnt(o, zsyn_tap(a.f, false, plo[i], phi[i], lo[i], hi[i]));
            if (z + 1 < a.az) ST::stnt(o + a.bx, zsyn_tap(a.f, true, plo[i], phi[i], lo[i], hi[i]));
          } else {
            if (z < a.az) ST::st(o, zsyn_tap(a.f, false, plo[i], phi[i], lo[i], hi[i]));
            if (z + 1 < a.az) ST::st(o + a.bx, zsyn_tap(a.f, true, plo[i], phi[i], lo[i], hi[i]));
          }
        }
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        plo[i] = lo[i];
        phi[i] = hi[i];
        an[i] = anx[i];
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) ann[i] = an[i];
  }
}

struct XYSynArgs {
  const void* B;
  long long bx, by, bvp;
  long long bzp, bry;  // z-plane and row pitch of B (bx by and bx, or bx and az bx)
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
// (tap 0 / 1), tile origins being even. AP: a.Ap is set (without it the LL
// band's next-level taps are compiled out).
template <class ST, int TX, bool AP>
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
  const E* src = (const E*)a.B + v * a.bvp + (long long)z * a.bzp;
  const E* ap = AP ? (const E*)a.Ap + v * a.apvp : nullptr;
  // the LL band of this z slice picks up syn_z(Ap, 0): taps j with z+2-j even and in range
  int za = -1, zb = -1;
  float wa = 0.f, wb = 0.f;
  if (AP) {
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
        lofs[q] = ((b >> 1) ? a.ncy + cy : cy) * (int)a.bry + ((b & 1) ? a.px : 0);
        if (b == 0 && AP) aofs[q] = cy * a.pa;
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
      if (AP && q < NAQ) {
        ba[q] = ST::zero();
        bb[q] = ST::zero();
      }
      if (lofs[q] >= 0 && cx < a.ncx) {
        bn[q] = src[lofs[q] + cx];  // (cached: neighbouring tiles share band rows and columns)
        if (AP && q < NAQ && aofs[q] >= 0) {
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
        if (AP && q < NAQ) {
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

// Per-level band-box geometry shared by the passes.
struct LevelBox {
  size_t px, bvol, pa, avol;  // band row half-width (ncx padded to 16), band box, next-input pitch and box
  int ncx, ncy, ncz;
  long long bx, by;
};

static std::vector<LevelBox> level_boxes(const std::vector<std::array<size_t, 3>>& in_dims,
                                         const std::vector<std::array<size_t, 3>>& cdims, int L) {
  std::vector<LevelBox> lb(L);
  for (int l = 0; l < L; ++l) {
    LevelBox& q = lb[l];
    q.ncx = (int)cdims[l][0] / 2;
    q.ncy = (int)cdims[l][1] / 2;
    q.ncz = (int)cdims[l][2] / 2;
    q.px = pad16(cdims[l][0] / 2);
    q.bx = 2 * (long long)q.px;
    q.by = (long long)cdims[l][1];
    q.bvol = 2 * q.px * cdims[l][1] * in_dims[l][2];
    q.pa = l > 0 ? pad16(in_dims[l][0]) : 0;
    q.avol = l > 0 ? q.pa * in_dims[l][1] * in_dims[l][2] : 0;
  }
  return lb;
}

// pass 1 of a level: xy analysis of nv volumes (or planes) `in` (row pitch ip1,
// ax x ay x az) into the band box B
template <class ST>
static void launch_xy_an(hipStream_t s, const void* in, long long ip1, const std::array<size_t, 3>& in_dims,
                         const LevelBox& q, void* B, size_t nv, const Db2F& f) {
  XYAnArgs a1{};
  a1.in = in;
  a1.ip1 = ip1;
  a1.ip2 = ip1 * (long long)in_dims[1];
  a1.ivp = a1.ip2 * (long long)in_dims[2];
  a1.B = B;
  a1.bx = q.bx;
  a1.by = q.by;
  a1.bvp = (long long)q.bvol;
  a1.px = (int)q.px;
  a1.ax = (int)in_dims[0];
  a1.ay = (int)in_dims[1];
  a1.az = (int)in_dims[2];
  a1.ncx = q.ncx;
  a1.ncy = q.ncy;
  a1.tiles_x = (q.ncx + AN_TCX - 1) / AN_TCX;
  a1.groups_x = (a1.tiles_x + TPW - 1) / TPW;
  a1.f = f;
  const int tiles_y = (q.ncy + AN_TCY - 1) / AN_TCY;
  dim3 g1((unsigned)(a1.groups_x * tiles_y), (unsigned)(in_dims[2] * nv));
  hipLaunchKernelGGL(k_wav_xy_an<ST>, g1, dim3(256), 0, s, a1);
}

// pass 3 of a level: xy synthesis of the band box B (plus syn_z(Ap, 0) on the
// LL band when Ap is set) into `out` (row pitch op1, the level's input box)
template <class ST>
static void launch_xy_syn(hipStream_t s, const void* B, const LevelBox& q, const void* Ap, long long apvp, int pa,
                          void* out, long long op1, const std::array<size_t, 3>& od, size_t nv, const Db2F& f,
                          bool zmid = false, int tpw = 0) {
  XYSynArgs a3{};
  a3.B = B;
  a3.bx = q.bx;
  a3.by = q.by;
  a3.bzp = zmid ? q.bx : q.bx * q.by;
  a3.bry = zmid ? q.bx * (long long)od[2] : q.bx;
  a3.bvp = (long long)q.bvol;
  a3.ncx = q.ncx;
  a3.ncy = q.ncy;
  a3.px = (int)q.px;
  a3.ncz = q.ncz;
  a3.Ap = Ap;
  a3.apvp = Ap ? apvp : 0;
  a3.pa = Ap ? pa : 0;
  a3.out = out;
  a3.op1 = op1;
  a3.op2 = a3.op1 * (long long)od[1];
  a3.ovp = a3.op2 * (long long)od[2];
  a3.ax = (int)od[0];
  a3.ay = (int)od[1];
  a3.az = (int)od[2];
  // 64-sample x tiles (128 for half storage measured slower), or 32 for rows
  // of at most 96 samples where they pad the row less (the deep levels' odd
  // lengths: 65 -> 96, not 128); 32-wide tiles of a 129-sample row take
  // 357 us against 231 us for 64-wide at 256x256x128x67's level 1
  const size_t ax = od[0];
  const bool narrow = ax <= 96 && (ax + 31) / 32 * 32 < (ax + SY_TX - 1) / SY_TX * SY_TX;
  const int TXS = narrow ? 32 : SY_TX;
  a3.tiles_x = (int)((od[0] + TXS - 1) / TXS);
  a3.tpw = tpw > 0 ? tpw : TPW;
  a3.groups_x = (a3.tiles_x + a3.tpw - 1) / a3.tpw;
  a3.f = f;
  const int tiles_y = (int)((od[1] + SY_TY - 1) / SY_TY);
  dim3 g3((unsigned)(a3.groups_x * tiles_y), (unsigned)(od[2] * nv));
  if (narrow) {
    if (Ap)
      hipLaunchKernelGGL((k_wav_xy_syn<ST, 32, true>), g3, dim3(256), 0, s, a3);
    else
      hipLaunchKernelGGL((k_wav_xy_syn<ST, 32, false>), g3, dim3(256), 0, s, a3);
  } else if (Ap) {
    hipLaunchKernelGGL((k_wav_xy_syn<ST, SY_TX, true>), g3, dim3(256), 0, s, a3);
  } else {
    hipLaunchKernelGGL((k_wav_xy_syn<ST, SY_TX, false>), g3, dim3(256), 0, s, a3);
  }
}

template <class ST>
static llr_status wavelet_run(llr_ctx* ctx, void* vols_v, size_t nvols, const size_t* dims, float alpha, int L,
                              void* coef_v) {
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
  const std::vector<LevelBox> lb = level_boxes(in_dims, cdims, L);
  std::vector<E*> Bp(L), Ap(L, nullptr);
  {
    E* p = (E*)coef_v;
    for (int l = 0; l < L; ++l) {
      Bp[l] = p;
      p += lb[l].bvol * nvols;
    }
    for (int l = 1; l < L; ++l) {
      Ap[l] = p;
      p += lb[l].avol * nvols;
    }
  }
  hipStream_t s = ctx->stream;
  for (int l = 0; l < L; ++l) {
    const size_t* in = in_dims[l].data();
    const LevelBox& q = lb[l];
    {
      KernelTimer kt(ctx, l == 0 ? kStageWavIn : kStageWavDeep);
      launch_xy_an<ST>(s, l == 0 ? (const void*)vols : (const void*)Ap[l], l == 0 ? (long long)in[0] : (long long)q.pa,
                       in_dims[l], q, Bp[l], nvols, f);
    }
    LLR_HIP(hipGetLastError());
    // pass 2
    ZArgs a2{};
    a2.B = Bp[l];
    a2.bx = q.bx;
    a2.by = q.by;
    a2.bvp = (long long)q.bvol;
    a2.az = (int)in[2];
    a2.ncz = q.ncz;
    a2.ncx = q.ncx;
    a2.ncy = q.ncy;
    a2.px = (int)q.px;
    a2.tiles_x = (int)((q.bx + ZT - 1) / ZT);
    a2.groups_x = (a2.tiles_x + TPW - 1) / TPW;
    a2.deepest = l == L - 1;
    a2.A = a2.deepest ? nullptr : Ap[l + 1];
    a2.avp = a2.deepest ? 0 : (long long)lb[l + 1].avol;
    a2.pa = a2.deepest ? 0 : (int)lb[l + 1].pa;
    a2.alpha = alpha;
    a2.f = f;
    size_t smem = (size_t)(in[2] + 2 * q.ncz) * ZT * sizeof(float2);  // LDS holds FP32 values
    if (smem > 160 * 1024 || in[2] > 512) return fail(LLR_ERR_UNSUPPORTED, "z length too large for the wavelet z pass");
    int nzr = (int)((in[2] + (256 / ZT) - 1) / (256 / ZT));
    dim3 g2((unsigned)(a2.groups_x * q.by), (unsigned)nvols);
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
    const LevelBox& q = lb[l];
    {
      KernelTimer kt(ctx, l == 0 ? kStageWavOut : kStageWavDeep);
      launch_xy_syn<ST>(s, Bp[l], q, l < L - 1 ? Ap[l + 1] : nullptr, l < L - 1 ? (long long)lb[l + 1].avol : 0,
                        l < L - 1 ? (int)lb[l + 1].pa : 0, l == 0 ? (void*)vols : (void*)Ap[l],
                        l == 0 ? (long long)in_dims[0][0] : (long long)q.pa, in_dims[l], nvols, f);
    }
    LLR_HIP(hipGetLastError());
  }
  return LLR_OK;
}

// K-plane workspace of the subspace wavelet (FP32 values): the z coefficients
// of every level and the next-level inputs of levels 1..L-1, each for K planes.
struct SubspaceWork {
  std::vector<float2*> WK, AK;
  std::vector<size_t> wvol;
};

static size_t subspace_work(const std::vector<std::array<size_t, 3>>& in_dims, const std::vector<LevelBox>& lb, int L,
                            int K, float2* base, SubspaceWork* w) {
  size_t off = 0;
  if (w) {
    w->WK.assign(L, nullptr);
    w->AK.assign(L, nullptr);
    w->wvol.assign(L, 0);
  }
  for (int l = 0; l < L; ++l) {
    const size_t wv = (size_t)lb[l].bx * (size_t)lb[l].by * 2 * (size_t)lb[l].ncz;
    if (w) {
      w->wvol[l] = wv;
      w->WK[l] = base + off;
    }
    off += wv * (size_t)K;
  }
  for (int l = 1; l < L; ++l) {
    if (w) w->AK[l] = base + off;
    off += lb[l].avol * (size_t)K;
  }
  (void)in_dims;
  return off;
}

template <class ST>
static llr_status wavelet_run_subspace(llr_ctx* ctx, void* vols_v, size_t nvols, const size_t* dims, float alpha,
                                       int L, void* coef_v, const WaveletFromSubspace& sub) {
  using E = typename ST::T;
  if (sub.K < 1 || sub.K > ZS_KM) return fail(LLR_ERR_UNSUPPORTED, "fused wavelet: rank <= 8");
  Db2F f = db2_filters();
  std::vector<std::array<size_t, 3>> in_dims, cdims;
  wavelet_geometry(dims, L, in_dims, cdims);
  const std::vector<LevelBox> lb = level_boxes(in_dims, cdims, L);
  SubspaceWork w;
  subspace_work(in_dims, lb, L, sub.K, sub.BK, &w);
  std::vector<E*> Bp(L), Ap(L, nullptr);
  {
    E* p = (E*)coef_v;
    for (int l = 0; l < L; ++l) {
      Bp[l] = p;
      p += lb[l].bvol * nvols;
    }
    for (int l = 1; l < L; ++l) {
      Ap[l] = p;
      p += lb[l].avol * nvols;
    }
  }
  hipStream_t s = ctx->stream;
  // analysis of the K planes, every level (one pass per level)
  for (int l = 0; l < L; ++l) {
    const LevelBox& q = lb[l];
    KernelTimer kt(ctx, l == 0 ? kStageWavIn : kStageWavDeep);
    An3Args z{};
    z.in = l == 0 ? sub.C : w.AK[l];
    z.ip1 = l == 0 ? (long long)in_dims[0][0] : (long long)q.pa;
    z.ip2 = z.ip1 * (long long)in_dims[l][1];
    z.ivp = z.ip2 * (long long)in_dims[l][2];
    z.WK = w.WK[l];
    z.wvp = (long long)w.wvol[l];
    const bool deepest = l == L - 1;
    z.AK = deepest ? nullptr : w.AK[l + 1];
    z.akvp = deepest ? 0 : (long long)lb[l + 1].avol;
    z.pa = deepest ? 0 : (int)lb[l + 1].pa;
    z.bx = q.bx;
    z.by = q.by;
    z.px = (int)q.px;
    z.ax = (int)in_dims[l][0];
    z.ay = (int)in_dims[l][1];
    z.az = (int)in_dims[l][2];
    z.ncx = q.ncx;
    z.ncy = q.ncy;
    z.ncz = q.ncz;
    z.tiles_x = (q.ncx + AN_TCX - 1) / AN_TCX;
    z.deepest = deepest;
    z.f = f;
    const unsigned tiles_y = (unsigned)((q.ncy + AN_TCY - 1) / AN_TCY);
    const dim3 g((unsigned)z.tiles_x * tiles_y, (unsigned)((q.ncz + AN3_NS - 1) / AN3_NS), (unsigned)sub.K);
    hipLaunchKernelGGL(k_wav_an3_k, g, dim3(AN3_TH), 0, s, z);
    LLR_HIP(hipGetLastError());
  }
  // synthesis of every volume, coarse to fine
  for (int l = L - 1; l >= 0; --l) {
    const LevelBox& q = lb[l];
    const int az = (int)in_dims[l][2];
    ZSynArgs a{};
    a.WK = w.WK[l];
    a.wvp = (long long)w.wvol[l];
    a.S = sub.S;
    a.K = sub.K;
    a.nq = (int)nvols;
    const bool deepest = l == L - 1;
    a.Ar = deepest ? nullptr : (const void*)Ap[l + 1];
    a.arvp = deepest ? 0 : (long long)lb[l + 1].avol;
    a.pa = deepest ? 0 : (int)lb[l + 1].pa;
    a.B = Bp[l];
    a.bvp = (long long)q.bvol;
    a.bx = q.bx;
    a.by = q.by;
    a.az = az;
    a.ncz = q.ncz;
    a.ncx = q.ncx;
    a.ncy = q.ncy;
    a.px = (int)q.px;
    a.deepest = deepest;
    a.alpha = alpha;
    a.f = f;
    {
      KernelTimer kt(ctx, l == 0 ? kStageWavMid : kStageWavDeep);
      const int nchunk = (az + 2 * ZM_CS - 1) / (2 * ZM_CS);
      const unsigned gx = (unsigned)(((q.bx + 31) / 32) * q.by * ((nchunk + 1) / 2));
      // volume ranges (whole groups of 16) so that a launch has about 8192
      // waves or more: the deep levels have few columns
      const int ngrp = (int)((nvols + 15) / 16);
      int nr = (int)((8192 + 4 * (size_t)gx - 1) / (4 * (size_t)gx));
      nr = nr < 1 ? 1 : (nr > ngrp ? ngrp : nr);
      const int qper = (ngrp + nr - 1) / nr * 16;
      const dim3 grid(gx, (unsigned)((nvols + qper - 1) / qper));
      // measured at 256x256x128x67's level 0: half storage 1.10 ms with the
      // workgroups of one column pair consecutive, 1.03 ms with column pairs
      // fastest (stores 1.20 ms non-temporal); complex64 1.51 ms, 1.40 ms with
      // non-temporal stores (also 1.85 -> 1.77 ms for the xy synthesis that
      // reads them), 1.69 ms with column pairs fastest
      constexpr bool h16 = ST::bytes == 4;
      if (sub.s_real)
        hipLaunchKernelGGL((k_wav_zsyn_m<ZM_CS, true, ST, !h16, h16>), grid, dim3(256), 0, s, a, sub.S, nchunk, qper);
      else
        hipLaunchKernelGGL((k_wav_zsyn_m<ZM_CS, false, ST, !h16, h16>), grid, dim3(256), 0, s, a, sub.S, nchunk, qper);
    }
    LLR_HIP(hipGetLastError());
    {
      KernelTimer kt(ctx, l == 0 ? kStageWavOut : kStageWavDeep);
      launch_xy_syn<ST>(s, Bp[l], q, nullptr, 0, 0, l == 0 ? vols_v : (void*)Ap[l],
                        l == 0 ? (long long)in_dims[0][0] : (long long)q.pa, in_dims[l], nvols, f, true, ctx->wav_tpw);
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
  return h16 ? wavelet_run<StoreF16>(ctx, vols, nvols, dims, alpha, L, coef)
              : wavelet_run<StoreF32>(ctx, vols, nvols, dims, alpha, L, coef);
}

size_t wavelet_subspace_elems(const size_t* dims, int L, int K) {
  if (L <= 0) return 0;
  std::vector<std::array<size_t, 3>> in_dims, cdims;
  wavelet_geometry(dims, L, in_dims, cdims);
  return subspace_work(in_dims, level_boxes(in_dims, cdims, L), L, K, nullptr, nullptr);
}

bool wavelet_subspace_supported(const size_t* dims, int L, int K) {
  (void)dims;
  return L >= 1 && K >= 1 && K <= ZS_KM;
}

llr_status wavelet_denoise_subspace(llr_ctx* ctx, void* vols, size_t nvols, const size_t* dims, float alpha, int L,
                                    void* coef, const WaveletFromSubspace& sub, bool h16) {
  if (!wavelet_subspace_supported(dims, L, sub.K)) return fail(LLR_ERR_UNSUPPORTED, "fused subspace wavelet");
  return h16 ? wavelet_run_subspace<StoreF16>(ctx, vols, nvols, dims, alpha, L, coef, sub)
             : wavelet_run_subspace<StoreF32>(ctx, vols, nvols, dims, alpha, L, coef, sub);
}

}  // namespace llr
