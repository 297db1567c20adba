// Mixed-precision form of the LLR step on gfx950 (LLR_PREC_MIXED).
//
//  - k_project_bf16 / k_backproject_bf16: the subspace projection C = X*S and
//    back-projection X = L*S^H (lib.rs:285-289) as real GEMMs on BF16 MFMA
//    (v_mfma_f32_16x16x16_bf16) with FP32 accumulation; S enters as two BF16
//    operands (hi + lo) so only the streamed X or L carries BF16 rounding
//    and the subspace itself is kept to ~2^-17. A 16-voxel tile is
//    the M dimension; the complex contraction is written as a real one over
//    interleaved (re, im) pairs, so K = 2*nq (projection) or 2*srank
//    (back-projection). Output columns are interleaved (re, im) too, so a
//    lane and its xor-1 neighbour hold the two parts of the same complex
//    value and each writes two voxels with one 16-byte store.
//  - k_llr_block_mixed: per shifted block, the Gram matrix G of the real
//    m x 2n matrix [Re F | Im F] on BF16 MFMA (both MFMA operands are the
//    same registers) with every value split exactly into three BF16 pieces
//    (Gram to ~2^-24, FP32 accumulation), and an FP32 Jacobi eigen-solve of
//    G giving V and the eigenvalues. If their condition number is above 1e6
//    the n x n eigen-solve is redone in FP64 (Gram of F in FP64 and a full
//    FP64 Jacobi). The block is replaced by F W, W = V diag(g) V^H.
//    (Earlier the Jacobi ran on the leading BF16 term only, followed by an
//    FP32 refinement sweep on V0^H G V0: the same accuracy, 0.2 ms more per
//    launch at 256x256x128 with 7 columns.)
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

#include "block_common.h"
#include "common.h"
#include "internal.h"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16 bit pattern, round to nearest even (one v_cvt_pk_bf16_f32 on gfx950)
__device__ __forceinline__ short bf16_bits(float f) { return __builtin_bit_cast(short, (__bf16)f); }

// D = A*B + C for 16x16 tiles, K = 16. Lane l supplies A[l%16][4(l/16)+j] and
// B[4(l/16)+j][l%16] (j = 0..3) and receives D[4(l/16)+j][l%16].
__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

constexpr int MAX_STEPS = 16;  // 2*nq <= 256

// BF16 value nearest to f, as float
__device__ __forceinline__ float bf16_round(float f) { return __uint_as_float(((unsigned)(unsigned short)bf16_bits(f)) << 16); }

// B fragments of the projection's real operand built from S (nq x K,
// column-major): row kk = 2q + ic, column n = 2k + oc, entries
// Re C += Re X Re S - Im X Im S, Im C += Re X Im S + Im X Re S, as BF16 hi + lo
// so the subspace the product uses is S to ~2^-17. S is read at clamped
// indices with no guard around the load, so a lane's loads are issued back to
// back and waited for once; entries outside the matrix are zeroed afterwards.
template <int K, int MS>
__device__ __forceinline__ void projection_operand(const float2* __restrict__ S, int nq, int n, int kq,
                                                   bf16x4 (&bop)[MS], bf16x4 (&bol)[MS]) {
  const int k = n >> 1, oc = n & 1;
  const int kc = k < K ? k : K - 1;
#pragma unroll
  for (int s = 0; s < MS; ++s) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int kk = 16 * s + 4 * kq + j, q = kk >> 1, ic = kk & 1;
      const float2 sv = S[(q < nq ? q : nq - 1) + (long long)kc * nq];
      float val = oc == 0 ? (ic == 0 ? sv.x : -sv.y) : (ic == 0 ? sv.y : sv.x);
      if (q >= nq || k >= K) val = 0.f;
      bop[s][j] = bf16_bits(val);
      bol[s][j] = bf16_bits(val - bf16_round(val));
    }
  }
}

// The same operand for FP16 MFMAs: S * 2^8 split into FP16 hi + lo (the scale
// keeps lo, about 2^-11 of S, in FP16's normal range for |S| >= 2^-8; the
// product is scaled back exactly by the caller). About 22 bits of S.
constexpr float P16_SSCALE = 256.f;
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
template <int K, int MS>
__device__ __forceinline__ void projection_operand_f16(const float2* __restrict__ S, int nq, int n, int kq,
                                                       f16x4 (&hop)[MS], f16x4 (&hol)[MS]) {
  const int k = n >> 1, oc = n & 1;
  const int kc = k < K ? k : K - 1;
#pragma unroll
  for (int s = 0; s < MS; ++s) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int kk = 16 * s + 4 * kq + j, q = kk >> 1, ic = kk & 1;
      const float2 sv = S[(q < nq ? q : nq - 1) + (long long)kc * nq];
      float val = oc == 0 ? (ic == 0 ? sv.x : -sv.y) : (ic == 0 ? sv.y : sv.x);
      if (q >= nq || k >= K) val = 0.f;
      val *= P16_SSCALE;
      const _Float16 h = (_Float16)val;
      hop[s][j] = h;
      hol[s][j] = (_Float16)(val - (float)h);
    }
  }
}

// Both projections work on tiles of PB_TV = 64 consecutive voxels (one
// 16-voxel MFMA row block per wavefront) and move the nq-plane side of the
// product through LDS, so every global access of a plane is a 512-byte run
// (64 voxels x 8 bytes) instead of a 128-byte one per wavefront.
constexpr int PB_TV = 64;
constexpr int PB_ROW = 2 * 8 * MAX_STEPS + 8;  // BF16 values per staged voxel row (padded)

// C[k][v] = sum_q X[q][v] S[q][k]. Row kk = 2q + c of the real operand is
// (Re, Im)[c] of X[q]; output column n = 2k + o is (Re, Im)[o] of C[k]. The
// X tile is staged in LDS as BF16 rows [voxel][Re X0, Im X0, Re X1, ...] so a
// lane's four K values are one 8-byte LDS read. MS bounds the K steps
// (2 nq <= 16 MS).
template <int K, int MS>
__global__ void __launch_bounds__(256) k_project_bf16(const float2* __restrict__ X, float2* __restrict__ Cp,
                                                      const float2* __restrict__ S, long long N, int nq) {
  __shared__ __align__(16) short s_x[PB_TV * PB_ROW];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int n = lane & 15, kq = lane >> 4;
  const int NS = (2 * nq + 15) / 16;
  // only the streamed data carries BF16 rounding
  bf16x4 bop[MS], bol[MS];
  projection_operand<K, MS>(S, nq, n, kq, bop, bol);
  const long long v0 = (long long)blockIdx.x * PB_TV;
  // stage: plane q, voxel vi -> row vi, columns 2q, 2q+1 (zeros past nq, up to
  // 16 NS); every load of the tile is issued before the first LDS store
  constexpr int NLD = 8 * MS * PB_TV / 256;
  // (clamped, unguarded loads: a guarded load was waited for on its own)
  float2 xl[NLD];
#pragma unroll
  for (int it = 0; it < NLD; ++it) {
    const int e = tid + 256 * it, vi = e % PB_TV, q = e / PB_TV;
    const long long vc = v0 + vi < N ? v0 + vi : N - 1;
    const float2 x = X[(long long)(q < nq ? q : nq - 1) * N + vc];
    xl[it] = (q < nq && v0 + vi < N) ? x : make_float2(0.f, 0.f);
  }
#pragma unroll
  for (int it = 0; it < NLD; ++it) {
    const int e = tid + 256 * it, vi = e % PB_TV, q = e / PB_TV;
    if (q < 8 * NS) {
      const unsigned u =
          (unsigned)(unsigned short)bf16_bits(xl[it].x) | ((unsigned)(unsigned short)bf16_bits(xl[it].y) << 16);
      *(unsigned*)&s_x[vi * PB_ROW + 2 * q] = u;
    }
  }
  __syncthreads();
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  const short* row = s_x + (16 * wave + n) * PB_ROW + 4 * kq;
#pragma unroll
  for (int s = 0; s < MS; ++s) {
    if (s < NS) {
      const uint2 u = *(const uint2*)(row + 16 * s);
      bf16x4 a = {(short)(u.x & 0xFFFF), (short)(u.x >> 16), (short)(u.y & 0xFFFF), (short)(u.y >> 16)};
      acc = mfma16(a, bop[s], acc);
      acc = mfma16(a, bol[s], acc);
    }
  }
  // lane holds column n of voxels 4kq..4kq+3; its xor-1 partner holds the other part
  const float p0 = __shfl_xor(acc[0], 1, 64), p1 = __shfl_xor(acc[1], 1, 64);
  const float p2 = __shfl_xor(acc[2], 1, 64), p3 = __shfl_xor(acc[3], 1, 64);
  const int k = n >> 1;
  if (k < K) {
    const long long w0 = v0 + 16 * wave + 4 * kq + ((n & 1) ? 2 : 0);
    float2 c0, c1;
    if ((n & 1) == 0) {
      c0 = make_float2(acc[0], p0);
      c1 = make_float2(acc[1], p1);
    } else {
      c0 = make_float2(p2, acc[2]);
      c1 = make_float2(p3, acc[3]);
    }
    float2* dst = Cp + (long long)k * N + w0;
    if (w0 + 1 < N) {
      dst[0] = c0;
      dst[1] = c1;
    } else if (w0 < N) {
      dst[0] = c0;
    }
  }
}

// X[q][v] = sum_k L[k][v] conj(S[q][k]) for voxels inside the block grid's
// coverage. Row kk = 2k + c is (Re, Im)[c] of L[k]; column n = 2(q - 8t) + o of
// output tile t is (Re, Im)[o] of X[q]. The product tile is gathered in LDS
// as [q][voxel] and written out plane by plane in 512-byte runs.
template <int K, int MS>
__global__ void __launch_bounds__(256) k_backproject_bf16(const float2* __restrict__ L, float2* __restrict__ X,
                                                          const float2* __restrict__ S, int nx, int ny, int nz,
                                                          int cx, int cy, int cz, int sx, int sy, int sz, int nq) {
  __shared__ __align__(16) float2 s_o[8 * MS * PB_TV];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int n = lane & 15, kq = lane >> 4;
  const int NT = (2 * nq + 15) / 16;
  bf16x4 bop[MS], bol[MS];  // S^H as hi + lo BF16 operands
#pragma unroll
  for (int t = 0; t < MS; ++t) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int kk = 4 * kq + j, k = kk >> 1, ic = kk & 1;
      const int nn = 16 * t + n, q = nn >> 1, oc = nn & 1;
      // clamped, unguarded load (issued with the others); zeroed outside S
      const float2 sv = S[(q < nq ? q : nq - 1) + (long long)(k < K ? k : K - 1) * nq];
      // Re X += Re L Re S + Im L Im S ; Im X += Im L Re S - Re L Im S
      float val = oc == 0 ? (ic == 0 ? sv.x : sv.y) : (ic == 0 ? -sv.y : sv.x);
      if (t >= NT || k >= K || q >= nq) val = 0.f;
      bop[t][j] = bf16_bits(val);
      bol[t][j] = bf16_bits(val - bf16_round(val));
    }
  }
  const long long N = (long long)nx * ny * nz;
  const long long v0 = (long long)blockIdx.x * PB_TV;
  const long long v = v0 + 16 * wave + n;
  const int k0 = 2 * kq;
  float2 l0 = make_float2(0.f, 0.f), l1 = l0;
  if (v < N) {
    if (k0 < K) l0 = L[(long long)k0 * N + v];
    if (k0 + 1 < K) l1 = L[(long long)(k0 + 1) * N + v];
  }
  const bf16x4 a = {bf16_bits(l0.x), bf16_bits(l0.y), bf16_bits(l1.x), bf16_bits(l1.y)};
#pragma unroll
  for (int t = 0; t < MS; ++t) {
    if (t < NT) {
      const f32x4 z4 = {0.f, 0.f, 0.f, 0.f};
      const f32x4 acc = mfma16(a, bol[t], mfma16(a, bop[t], z4));
      // D[4kq + j][n]: voxel 16 wave + 4kq + j, column n = (volume 8t + n/2, part n & 1)
      const int q = 8 * t + (n >> 1);
      float* o = (float*)(s_o + (long long)q * PB_TV + 16 * wave + 4 * kq) + (n & 1);
#pragma unroll
      for (int j = 0; j < 4; ++j) o[2 * j] = acc[j];
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < 8 * MS * PB_TV / 256; ++it) {
    const int e = tid + 256 * it, vi = e % PB_TV, q = e / PB_TV;
    const long long w = v0 + vi;
    if (q >= nq || w >= N) continue;
    const int x = (int)(w % nx), y = (int)((w / nx) % ny), z = (int)(w / ((long long)nx * ny));
    int rx = x - sx, ry = y - sy, rz = z - sz;
    if (rx < 0) rx += nx;
    if (ry < 0) ry += ny;
    if (rz < 0) rz += nz;
    if (rx < cx && ry < cy && rz < cz) X[(long long)q * N + w] = s_o[q * PB_TV + vi];
  }
}

__global__ void __launch_bounds__(256) k_llr_block_mixed(BlockArgs A) {
  extern __shared__ double2 smem_d[];  // 16-byte aligned base
  __shared__ int s_fp64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nc = A.nc;
  const int m = (int)A.g.m, mp = m + 1;  // m < 2^20 (geom_in_range)
  const int fsz = (nc * mp + 1) & ~1;     // keeps what follows 16-byte aligned
  const long long slot = xcd_slot(A.per_xcd);
  if (slot >= A.nblocks) return;
  const long long b = A.block0 + slot;
  float2* Fs = (float2*)smem_d;                    // nc * mp: the block F
  float* Rp = (float*)(Fs + fsz);                  // 4 x 256 floats: per-wave BF16 Gram partials (leading term)
  float* Rq = Rp + 1024;                           // 4 x 256 floats: the split Gram's remaining terms
  double2* Gp64 = (double2*)Rp;                    // 256 double2 (FP64 fallback, aliases Rp)
  float2* Ws = (float2*)(Rq + 1024);               // 64: W = J diag(g) J^H

  const BlockRows br(A.g, b);
  const long long cs = col_stride(A);
  // two rows per thread at a time, every load issued before any store (a
  // store between them would keep each load's latency in turn)
  const bool prev_in = A.prev && !A.mom_first;
  for (int i0 = tid; i0 < m; i0 += 512) {
    long long base[2];
    float2 v[2][8], pv[2][8];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int i = i0 + 256 * h;
      base[h] = i < m ? row_base(A, br, b, i) : -1;
#pragma unroll
      for (int col = 0; col < 8; ++col) {
        v[h][col] = make_float2(0.f, 0.f);
        pv[h][col] = make_float2(0.f, 0.f);
        if (col < nc && base[h] >= 0) {
          v[h][col] = A.F[base[h] + col * cs];
          if (prev_in) pv[h][col] = A.prev[base[h] + col * cs];
        }
      }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int i = i0 + 256 * h;
      if (i >= m) continue;
#pragma unroll
      for (int col = 0; col < 8; ++col) {
        if (col >= nc) continue;
        float2 x = v[h][col];
        if (A.prev && base[h] >= 0) {  // momentum: every coefficient belongs to exactly one block
          const float2 q = A.mom_first ? x : pv[h][col];
          A.prev[base[h] + col * cs] = x;
          x = make_float2((1.f + A.mom) * x.x - A.mom * q.x, (1.f + A.mom) * x.y - A.mom * q.y);
        }
        Fs[col * mp + i] = x;
      }
    }
  }
  __syncthreads();

  // 1. Gram of the real m x 2nc matrix [Re F | Im F] on BF16 MFMA (16 rows per
  //    MFMA). Every value is split exactly into three BF16 pieces x = h + u + l;
  //    the leading term h h^T is the BF16 Gram that seeds the eigenvectors, and
  //    with h u^T + u h^T + h l^T + l h^T + u u^T it is the Gram to ~2^-24 (the
  //    dropped u l^T, l u^T, l l^T are smaller), accumulated in FP32.
  {
    const int cidx = lane & 15, kq = lane >> 4;
    const bool valid = cidx < 2 * nc, re = cidx < nc;
    const float2* colp = Fs + (long long)(valid ? (re ? cidx : cidx - nc) : 0) * mp;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f}, acc3 = {0.f, 0.f, 0.f, 0.f};
    const int tiles = (m + 15) / 16;
    for (int t = wave; t < tiles; t += 4) {
      bf16x4 vh, vu, vl;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        int row = t * 16 + 4 * kq + j;
        float x = 0.f;
        if (valid && row < m) {
          float2 f = colp[row];
          x = re ? f.x : f.y;
        }
        const float h = bf16_round(x), r1 = x - h, u = bf16_round(r1);
        vh[j] = bf16_bits(h);
        vu[j] = bf16_bits(u);
        vl[j] = bf16_bits(r1 - u);
      }
      acc = mfma16(vh, vh, acc);
      acc3 = mfma16(vh, vu, acc3);
      acc3 = mfma16(vu, vh, acc3);
      acc3 = mfma16(vh, vl, acc3);
      acc3 = mfma16(vl, vh, acc3);
      acc3 = mfma16(vu, vu, acc3);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      Rp[wave * 256 + (4 * kq + j) * 16 + cidx] = acc[j];
      Rq[wave * 256 + (4 * kq + j) * 16 + cidx] = acc3[j];
    }
  }
  __syncthreads();

  // 2. wave 0: the complex Gram G = F^H F from the real one (the split
  //    Gram's terms summed), its FP32 Jacobi eigen-solve -> eigenvectors jv,
  //    eigenvalues on hv's diagonal; condition test
  const int r = lane & 7, c = lane >> 3;
  float2 hv = make_float2(0.f, 0.f), jv = make_float2(0.f, 0.f);
  if (wave == 0) {
    float2 g3 = make_float2(0.f, 0.f);
    if (r < nc && c < nc) {
      float rr = 0.f, ii = 0.f, ri = 0.f, ir = 0.f, rr3 = 0.f, ii3 = 0.f, ri3 = 0.f, ir3 = 0.f;
      for (int w = 0; w < 4; ++w) {
        const float* R = Rp + w * 256;
        const float* Q = Rq + w * 256;
        rr += R[r * 16 + c];
        ii += R[(nc + r) * 16 + nc + c];
        ri += R[r * 16 + nc + c];
        ir += R[(nc + r) * 16 + c];
        rr3 += Q[r * 16 + c];
        ii3 += Q[(nc + r) * 16 + nc + c];
        ri3 += Q[r * 16 + nc + c];
        ir3 += Q[(nc + r) * 16 + c];
      }
      g3 = make_float2((rr + rr3) + (ii + ii3), r == c ? 0.f : (ri + ri3) - (ir + ir3));
    }
    wave_jacobi8<float>(g3, jv, lane, 16, 2.5e-7f);
    hv = g3;
    float lmax = 0.f, lmin = FLT_MAX;
#pragma unroll
    for (int l = 0; l < 8; ++l) {
      float x = __shfl(hv.x, 9 * l, 64);
      if (l < nc) {
        lmax = fmaxf(lmax, x);
        lmin = fminf(lmin, x);
      }
    }
    if (lane == 0) s_fp64 = (lmax > 0.f && !(lmin > 1e-6f * lmax)) ? 1 : 0;
  }
  __syncthreads();

  // 3. FP64 fallback for ill-conditioned blocks: Gram of F and Jacobi in FP64;
  //    its eigenvectors replace V0 J
  const bool fb = s_fp64 != 0;
  if (fb) {
    const int NE = nc * (nc + 1) / 2;
    const int NH = 256 / NE;
    int gi = 0, gj = 0;
    {
      int rem = tid % NE;
      for (int jj = 0; jj < nc; ++jj) {
        if (rem <= jj) {
          gi = rem;
          gj = jj;
          break;
        }
        rem -= jj + 1;
      }
      const int hh = tid / NE;
      double2 acc = make_double2(0.0, 0.0);
      if (hh < NH)
        for (int row = hh; row < m; row += NH) {
          float2 x = Fs[gi * mp + row], y = Fs[gj * mp + row];
          double2 xd = make_double2(x.x, x.y), yd = make_double2(y.x, y.y);
          acc = cadd(acc, cconjmul(xd, yd));
        }
      Gp64[tid] = acc;
    }
    __syncthreads();
    if (wave == 0) {
      double2 a64 = make_double2(0.0, 0.0), v64;
      if (r < nc && c < nc) {
        int i = min(r, c), j = max(r, c);
        int e = j * (j + 1) / 2 + i;
        double2 sd = make_double2(0.0, 0.0);
        for (int hh = 0; hh < NH; ++hh) sd = cadd(sd, Gp64[hh * NE + e]);
        a64 = (r <= c) ? sd : make_double2(sd.x, -sd.y);
        if (r == c) a64.y = 0.0;
      }
      wave_jacobi8<double>(a64, v64, lane, 30, 1e-15);
      hv = make_float2((float)a64.x, 0.f);
      jv = make_float2((float)v64.x, (float)v64.y);
      if (lane == 0 && A.fp64_count) atomicAdd(A.fp64_count, 1u);
    }
  }

  // 4. wave 0: weights, W = J diag(g) J^H (J the FP32 or the FP64 eigenvectors)
  if (wave == 0) {
    float lam[8];
#pragma unroll
    for (int l = 0; l < 8; ++l) lam[l] = __shfl(hv.x, 9 * l, 64);
    float lmax = 0.f;
#pragma unroll
    for (int l = 0; l < 8; ++l)
      if (l < nc) lmax = fmaxf(lmax, lam[l]);
    if (lane == 0 && A.smax_bits) atomicMax(A.smax_bits, __float_as_uint(sqrtf(fmaxf(lmax, 0.f))));
    if (A.svals && lane < 8) A.svals[slot * 8 + lane] = sqrtf(fmaxf(lam[lane], 0.f));
    if (A.mode != 2) {
      float2 w = make_float2(0.f, 0.f);  // J diag(g) J^H
#pragma unroll
      for (int l = 0; l < 8; ++l) {
        float g = shrink_weight(lam, l, nc, A.mode, A.rank, A.lambda);
        float2 t = cmulc(shfl2(jv, r + 8 * l), shfl2(jv, c + 8 * l));
        w.x += g * t.x;
        w.y += g * t.y;
      }
      Ws[lane] = w;  // W[r][c]
    }
  }
  if (A.mode == 2) return;
  __syncthreads();

  // 5. out = F W (rows outside a partial edge block are not written)
  for (int rr = tid; rr < m; rr += 256) {
    const long long base = row_base(A, br, b, rr);
    if (base < 0) continue;
    float2 f[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) f[k] = k < nc ? Fs[k * mp + rr] : make_float2(0.f, 0.f);
    for (int cc = 0; cc < nc; ++cc) {
      float2 acc = make_float2(0.f, 0.f);
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (k < nc) acc = cadd(acc, cmul(f[k], Ws[k + 8 * cc]));
      A.out[base + cc * cs] = acc;
    }
  }
}

// Projection C = X*S of an iterate held in half storage, on FP16 MFMA with FP32
// accumulation: the stored half pairs are the A operand as they are (no
// conversion), and S (times 2^8) enters as FP16 hi + lo, two MFMAs per K step,
// so it adds no rounding beyond the storage itself and S's ~2^-22. (Taking X
// as the BF16 pair hi + lo instead, hi*S_hi + hi*S_lo + lo*S_hi on BF16 MFMA,
// cost three MFMAs and about 20 conversion instructions per step: 0.78 ms per
// launch at 256x256x128x67.) Tiles of P16_TV = 128 voxels (two 16-voxel MFMA
// row blocks per wavefront); plane runs of 512 bytes are staged in LDS as rows
// [voxel][plane] of 32-bit half pairs.
constexpr int P16_TV = 128;
constexpr int P16_TPW = 8;  // tiles per workgroup (large volumes; at 256x256x128x67, swept over 1-16: 0.638 ms at 8, 0.686 at 4)

template <int K, int MS>
__global__ void __launch_bounds__(256) k_project_s16(const ch16* __restrict__ X, float2* __restrict__ Cp,
                                                     const float2* __restrict__ S, long long N, int nq,
                                                     float out_scale, int tpw) {
  constexpr int ROW = 8 * MS + 2;  // 32-bit words per staged voxel row (even: 8-byte aligned reads)
  __shared__ __align__(16) unsigned s_x[P16_TV * ROW];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int n = lane & 15, kq = lane >> 4;
  const int NS = (2 * nq + 15) / 16;
  f16x4 hop[MS], hol[MS];
  projection_operand_f16<K, MS>(S, nq, n, kq, hop, hol);
  const float oscale = out_scale / P16_SSCALE;
  // tpw consecutive tiles per workgroup; the next tile's words are loaded
  // into registers while the current one is multiplied
  const long long tiles = (N + P16_TV - 1) / P16_TV;
  const long long t0 = (long long)blockIdx.x * tpw;
  const int nt = (int)min((long long)tpw, tiles - t0);
  constexpr int NLD = 8 * MS * P16_TV / 256;
  unsigned xl[NLD];
  auto load = [&](long long v0) {
#pragma unroll
    for (int it = 0; it < NLD; ++it) {
      const int e = tid + 256 * it, vi = e % P16_TV, q = e / P16_TV;
      xl[it] = 0u;
      if (q < nq && v0 + vi < N) xl[it] = StoreF16::rawnt(X + (long long)q * N + v0 + vi);
    }
  };
  load(t0 * P16_TV);
  for (int ti = 0; ti < nt; ++ti) {
    const long long v0 = (t0 + ti) * P16_TV;
    if (ti > 0) __syncthreads();  // the previous tile's LDS reads are done
#pragma unroll
    for (int it = 0; it < NLD; ++it) {
      const int e = tid + 256 * it, vi = e % P16_TV, q = e / P16_TV;
      if (q < 8 * NS) s_x[vi * ROW + q] = xl[it];
    }
    __syncthreads();
    if (ti + 1 < nt) load(v0 + P16_TV);
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      const unsigned* row = s_x + (32 * wave + 16 * rb + n) * ROW + 2 * kq;
#pragma unroll
      for (int s = 0; s < MS; ++s) {
        if (s < NS) {
          // (Re X_q, Im X_q, Re X_q+1, Im X_q+1): the lane's four K values
          const f16x4 a = __builtin_bit_cast(f16x4, *(const uint2*)(row + 8 * s));
          acc = __builtin_amdgcn_mfma_f32_16x16x16f16(a, hop[s], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x16f16(a, hol[s], acc, 0, 0, 0);
        }
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[j] *= oscale;
      const float p0 = __shfl_xor(acc[0], 1, 64), p1 = __shfl_xor(acc[1], 1, 64);
      const float p2 = __shfl_xor(acc[2], 1, 64), p3 = __shfl_xor(acc[3], 1, 64);
      const int k = n >> 1;
      if (k < K) {
        const long long w0 = v0 + 32 * wave + 16 * rb + 4 * kq + ((n & 1) ? 2 : 0);
        float2 c0, c1;
        if ((n & 1) == 0) {
          c0 = make_float2(acc[0], p0);
          c1 = make_float2(acc[1], p1);
        } else {
          c0 = make_float2(p2, acc[2]);
          c1 = make_float2(p3, acc[3]);
        }
        float2* dst = Cp + (long long)k * N + w0;
        if (w0 + 1 < N) {
          dst[0] = c0;
          dst[1] = c1;
        } else if (w0 < N) {
          dst[0] = c0;
        }
      }
    }
  }
}

}  // namespace

namespace llr {

llr_status project_s16_device(llr_ctx* ctx, const void* X, float2* Cp, const float2* S, size_t N, int nq, int K,
                              float out_scale) {
  if (nq > 8 * MAX_STEPS) return fail(LLR_ERR_UNSUPPORTED, "BF16 projection handles at most 128 volumes");
  // 128-voxel tiles, P16_TPW per workgroup while that leaves at least 4096 workgroups
  const size_t tiles = (N + P16_TV - 1) / P16_TV;
  int tpw = P16_TPW;
  while (tpw > 1 && (tiles + tpw - 1) / tpw < 4096) tpw /= 2;
  if (ctx->proj_tpw > 0) tpw = ctx->proj_tpw;
  const unsigned grid = (unsigned)((tiles + tpw - 1) / tpw);
  const bool small = 2 * nq <= 16 * 9;
  const ch16* Xs = (const ch16*)X;
#define PS_CASE(KK)                                                                                               \
  case KK:                                                                                                        \
    if (small)                                                                                                    \
      hipLaunchKernelGGL((k_project_s16<KK, 9>), dim3(grid), dim3(256), 0, ctx->stream, Xs, Cp, S, (long long)N, nq, \
                         out_scale, tpw);                                                                         \
    else                                                                                                          \
      hipLaunchKernelGGL((k_project_s16<KK, MAX_STEPS>), dim3(grid), dim3(256), 0, ctx->stream, Xs, Cp, S,          \
                         (long long)N, nq, out_scale, tpw);                                                       \
    break;
  switch (K) {
    PS_CASE(1) PS_CASE(2) PS_CASE(3) PS_CASE(4) PS_CASE(5) PS_CASE(6) PS_CASE(7) PS_CASE(8)
    default: return fail(LLR_ERR_UNSUPPORTED, "BF16 projection handles subspace ranks up to 8");
  }
#undef PS_CASE
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

llr_status project_bf16_device(llr_ctx* ctx, const float2* X, float2* Cp, const float2* S, size_t N, int nq, int K) {
  if (nq > 8 * MAX_STEPS) return fail(LLR_ERR_UNSUPPORTED, "BF16 projection handles at most 128 volumes");
  const unsigned grid = (unsigned)((N + PB_TV - 1) / PB_TV);
  const bool small = 2 * nq <= 16 * 9;
#define PB_CASE(KK)                                                                                              \
  case KK:                                                                                                       \
    if (small)                                                                                                   \
      hipLaunchKernelGGL((k_project_bf16<KK, 9>), dim3(grid), dim3(256), 0, ctx->stream, X, Cp, S, (long long)N, nq); \
    else                                                                                                         \
      hipLaunchKernelGGL((k_project_bf16<KK, MAX_STEPS>), dim3(grid), dim3(256), 0, ctx->stream, X, Cp, S,          \
                         (long long)N, nq);                                                                      \
    break;
  switch (K) {
    PB_CASE(1) PB_CASE(2) PB_CASE(3) PB_CASE(4) PB_CASE(5) PB_CASE(6) PB_CASE(7) PB_CASE(8)
    default: return fail(LLR_ERR_UNSUPPORTED, "BF16 projection handles subspace ranks up to 8");
  }
#undef PB_CASE
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

llr_status backproject_bf16_device(llr_ctx* ctx, const float2* C, float2* X, const float2* S, const size_t* vol,
                                   const size_t* blk, const size_t* shift, int nq, int K, bool partial) {
  size_t N = vol[0] * vol[1] * vol[2];
  if (N >= 0xFFFFFFFFull) return fail(LLR_ERR_UNSUPPORTED, "volume larger than 2^32 voxels");
  if (nq > 8 * MAX_STEPS) return fail(LLR_ERR_UNSUPPORTED, "BF16 back-projection handles at most 128 volumes");
  int cov[3], sh[3];
  for (int a = 0; a < 3; ++a) {
    cov[a] = (int)(partial ? vol[a] : (vol[a] / blk[a]) * blk[a]);
    sh[a] = (int)(shift[a] % vol[a]);
  }
  const unsigned grid = (unsigned)((N + PB_TV - 1) / PB_TV);
  const bool small = 2 * nq <= 16 * 9;
#define BB_CASE(KK)                                                                                              \
  case KK:                                                                                                       \
    if (small)                                                                                                   \
      hipLaunchKernelGGL((k_backproject_bf16<KK, 9>), dim3(grid), dim3(256), 0, ctx->stream, C, X, S, (int)vol[0], \
                         (int)vol[1], (int)vol[2], cov[0], cov[1], cov[2], sh[0], sh[1], sh[2], nq);             \
    else                                                                                                         \
      hipLaunchKernelGGL((k_backproject_bf16<KK, MAX_STEPS>), dim3(grid), dim3(256), 0, ctx->stream, C, X, S,     \
                         (int)vol[0], (int)vol[1], (int)vol[2], cov[0], cov[1], cov[2], sh[0], sh[1], sh[2], nq); \
    break;
  switch (K) {
    BB_CASE(1) BB_CASE(2) BB_CASE(3) BB_CASE(4) BB_CASE(5) BB_CASE(6) BB_CASE(7) BB_CASE(8)
    default: return fail(LLR_ERR_UNSUPPORTED, "BF16 back-projection handles subspace ranks up to 8");
  }
#undef BB_CASE
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

size_t llr_block_mixed_smem(long long m, int nc) {
  size_t fsz = ((size_t)nc * (size_t)(m + 1) + 1) & ~(size_t)1;
  return (fsz + 512 + 512 + 64) * sizeof(float2);
}

llr_status llr_block_run_mixed(llr_ctx* ctx, float2* F, const size_t* vol, const size_t* blk, const size_t* shift,
                               int nc, int layout, long long block0, long long nblocks, int mode, int rank,
                               float lambda, unsigned* smax_bits, float* svals, unsigned* fp64_count, bool partial,
                               float2* prev, float mom, bool mom_first) {
  GeomArgs g = make_geom(vol, blk, shift, partial);
  if (nc < 1 || nc > 8) return fail(LLR_ERR_UNSUPPORTED, "mixed LLR kernel handles 1..8 columns");
  if (!geom_in_range(g)) return fail(LLR_ERR_UNSUPPORTED, "block geometry beyond the LLR kernel's index range");
  if (nblocks <= 0) return LLR_OK;
  size_t smem = llr_block_mixed_smem(g.m, nc);
  if (smem > 64 * 1024) return fail(LLR_ERR_UNSUPPORTED, "block too large for the mixed LLR kernel");
  BlockArgs A{};
  A.F = F;
  A.out = F;
  A.S = nullptr;
  A.g = g;
  A.nc = nc;
  A.nq = 0;
  A.layout = layout;
  A.nblocks = nblocks;
  A.block0 = block0;
  A.per_xcd = (int)((nblocks + 7) / 8);
  A.mode = mode;
  A.rank = rank;
  A.lambda = lambda;
  A.smax_bits = smax_bits;
  A.svals = svals;
  A.stage_lds = 1;
  A.fp64_count = fp64_count;
  if (prev && layout != 0) return fail(LLR_ERR_ARG, "momentum on load needs the voxel-plane layout");
  A.prev = prev;
  A.mom = mom;
  A.mom_first = mom_first ? 1 : 0;
  hipLaunchKernelGGL(k_llr_block_mixed, dim3((unsigned)(8 * A.per_xcd)), dim3(256), smem, ctx->stream, A);
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

}  // namespace llr
