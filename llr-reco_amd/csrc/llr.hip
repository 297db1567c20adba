// LLR proximal step on gfx950.
//
//  - k_lift / k_unlift: the shifted-block gather/scatter of block.rs:8-184 as
//    pure data movement (bit-exact), any element size.
//  - k_project: C[k][v] = sum_q X[q][v] S[q][k], the subspace projection X*S of
//    dti_project + conj_transpose (lib.rs:285-286) done per voxel over the
//    volume-major iterate, fully coalesced.
//  - k_llr_block: one workgroup per shifted block: gather the m x n Casorati
//    matrix F (n <= 8) into LDS, Gram G = F^H F, Hermitian Jacobi eigen-solve of
//    G by one wavefront (one matrix element per lane, the 7-round round-robin
//    ordering of 8 indices), W = V diag(g(s)) V^H, then F*W (or F*W*S^H with a
//    subspace, lib.rs:287-289) scattered back. g(s) = 1 for the top `rank`
//    singular values (low_rank.rs:10-69) or max(s - lambda, 0)/s (SVT).
//  - generic kernels for matrices with n > 8 columns (host drop-in API only):
//    Gram, serial per-matrix Jacobi, apply.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>

#include "block_common.h"
#include "common.h"
#include "internal.h"

namespace {


// Gather (lift, block.rs:147-164) or scatter (un_lift, block.rs:166-184) of
// whole blocks, by rows of the block grid. Workgroup (row, y, z) owns the
// grid_x blocks of one grid row (the matrices c = row grid_x + bx that fall in
// [start, start + n)), the volumes q = 4y .. 4y+3 and a run of block rows
// r = iy + blk[1] iz (blockIdx.z). Thread t owns the voxel column x = t,
// t + 256, ... of the row, so every volume row the workgroup touches is one
// contiguous run of grid_x blk[0] voxels, and a thread's matrix offsets
// advance by blk[0] per block row (a block's matrix fills contiguously in
// L2). Four block rows of four volumes are loaded before the first value is
// stored. Matrix c, volume q, row i is at ((c - start) nvols + q) m + i.
template <class E, bool UNLIFT>
__global__ void __launch_bounds__(256) k_lift_rows(const E* __restrict__ src, E* __restrict__ dst, GeomArgs g,
                                                    long long nvols, long long start, long long n) {
  constexpr int QU = 4, RU = 4;
  const int gx = (int)g.grid[0], gyy = (int)g.grid[1];
  const int b0 = (int)g.blk[0], b1 = (int)g.blk[1], b2 = (int)g.blk[2], m = (int)g.m;
  const int nx = (int)g.vol[0], ny = (int)g.vol[1], nz = (int)g.vol[2];
  const long long row = blockIdx.x + start / gx;  // grid row (by, bz)
  const int by = (int)(row % gyy), bz = (int)(row / gyy);
  const int nr = b1 * b2;
  // this workgroup's block rows [rlo, rhi): grid.z splits them in runs of RU
  const int rper = ((nr + (int)gridDim.z - 1) / (int)gridDim.z + RU - 1) / RU * RU;
  const int rlo = blockIdx.z * rper, rhi = min(nr, rlo + rper);
  // rows may be wider than small_div's exact range (2^20): an integer divide,
  // once per voxel column
  for (int xl = threadIdx.x; xl < gx * b0; xl += blockDim.x) {
    const int bx = xl / b0, ix = xl - bx * b0;
    const long long c = row * gx + bx;
    if (c < start || c >= start + n) continue;
    int x = bx * b0 + ix + (int)g.shift[0];
    if (x >= nx) x -= nx;
    const long long mbase = (c - start) * nvols * m + ix;
    for (long long q0 = (long long)blockIdx.y * QU; q0 < nvols; q0 += (long long)gridDim.y * QU) {
      for (int r0 = rlo; r0 < rhi; r0 += RU) {
        E val[RU][QU];
        long long va[RU];
#pragma unroll
        for (int rr = 0; rr < RU; ++rr) {
          const int r = r0 + rr;
          va[rr] = -1;
          if (r < rhi) {
            const int iz = r / b1, iy = r - iz * b1;
            int y = by * b1 + iy + (int)g.shift[1], z = bz * b2 + iz + (int)g.shift[2];
            if (y >= ny) y -= ny;
            if (z >= nz) z -= nz;
            va[rr] = x + (long long)nx * (y + (long long)ny * z);
          }
        }
#pragma unroll
        for (int rr = 0; rr < RU; ++rr)
#pragma unroll
          for (int u = 0; u < QU; ++u) {
            const long long q = q0 + u;
            if (q < nvols && va[rr] >= 0)
              val[rr][u] = UNLIFT ? src[mbase + q * m + (long long)(r0 + rr) * b0] : src[q * g.N + va[rr]];
          }
#pragma unroll
        for (int rr = 0; rr < RU; ++rr)
#pragma unroll
          for (int u = 0; u < QU; ++u) {
            const long long q = q0 + u;
            if (q < nvols && va[rr] >= 0) {
              if (UNLIFT)
                dst[q * g.N + va[rr]] = val[rr][u];
              else
                dst[mbase + q * m + (long long)(r0 + rr) * b0] = val[rr][u];
            }
          }
      }
    }
  }
}

// non-temporal load / store of a 4-, 8- or 16-byte element
template <class E>
__device__ __forceinline__ E nt_load(const E* p) {
  typedef unsigned int v2u __attribute__((ext_vector_type(2)));
  typedef unsigned int v4u __attribute__((ext_vector_type(4)));
  if constexpr (sizeof(E) == 4) return __builtin_bit_cast(E, __builtin_nontemporal_load((const unsigned int*)p));
  else if constexpr (sizeof(E) == 8) return __builtin_bit_cast(E, __builtin_nontemporal_load((const v2u*)p));
  else return __builtin_bit_cast(E, __builtin_nontemporal_load((const v4u*)p));
}
template <class E>
__device__ __forceinline__ void nt_store(const E& v, E* p) {
  typedef unsigned int v2u __attribute__((ext_vector_type(2)));
  typedef unsigned int v4u __attribute__((ext_vector_type(4)));
  if constexpr (sizeof(E) == 4) __builtin_nontemporal_store(__builtin_bit_cast(unsigned int, v), (unsigned int*)p);
  else if constexpr (sizeof(E) == 8) __builtin_nontemporal_store(__builtin_bit_cast(v2u, v), (v2u*)p);
  else __builtin_nontemporal_store(__builtin_bit_cast(v4u, v), (v4u*)p);
}

// Gather/scatter through LDS, one slab at a time. Workgroup (x, y, z) owns
// the blocks [xc xbw, xc xbw + xbw) of grid row `row` (blockIdx.x = row nxc +
// xc), the block rows r0 .. r0 + RC - 1 (r0 = RC blockIdx.z) and the volumes
// q = y, y + gridDim.y, ... For one volume the slab is RC volume runs of
// xbw blk[0] voxels; for each of its blocks the same slab is RC blk[0]
// consecutive matrix elements starting at row r0 blk[0]. Lift reads the
// volume runs (coalesced along x) into the LDS tile and writes each block's
// run contiguously; un_lift does the reverse. Tile rows are padded by 8
// elements, so the 8 x 8 elements a wavefront touches on the block side fall
// on distinct LDS banks. Matrix c, volume q, row i is at
// ((c - start) nvols + q) m + i (block.rs:147-184). NT bit 0: non-temporal
// loads, bit 1: non-temporal stores.
constexpr int LB_PAD = 8;
template <class E, bool UNLIFT, int RC, int NT>
__global__ void __launch_bounds__(256) k_lift_slab(const E* __restrict__ src, E* __restrict__ dst, GeomArgs g,
                                                    long long nvols, long long start, long long n, int nxc, int xbw) {
  static_assert(RC <= 256, "row offsets are filled by one thread each");
  extern __shared__ unsigned char lift_lds[];
  E* tile = reinterpret_cast<E*>(lift_lds);
  __shared__ long long rowoff[RC];
  constexpr int NB = 8;  // elements per thread loaded before the first store
  const int gx = (int)g.grid[0], gyy = (int)g.grid[1];
  const int b0 = (int)g.blk[0], b1 = (int)g.blk[1], m = (int)g.m;
  const int nx = (int)g.vol[0], ny = (int)g.vol[1], nz = (int)g.vol[2];
  const int WP = min(gx, xbw) * b0 + LB_PAD;
  const long long rowi = blockIdx.x / nxc;
  const int xc = (int)(blockIdx.x - rowi * nxc);
  const long long row = rowi + start / gx;
  const long long c0 = row * gx;
  const int by = (int)(row % gyy), bz = (int)(row / gyy);
  const int r0 = (int)blockIdx.z * RC;
  const int rc = min(RC, b1 * (int)g.blk[2] - r0);
  const int bxlo = (int)max((long long)xc * xbw, start - c0);
  const int bxhi = (int)min((long long)min(gx, xc * xbw + xbw), start + n - c0);
  if (bxlo >= bxhi) return;  // the whole workgroup, before any barrier
  if (threadIdx.x < rc) {
    const int r = r0 + (int)threadIdx.x, iz = r / b1, iy = r - iz * b1;
    int y = by * b1 + iy + (int)g.shift[1], z = bz * (int)g.blk[2] + iz + (int)g.shift[2];
    if (y >= ny) y -= ny;
    if (z >= nz) z -= nz;
    rowoff[threadIdx.x] = (long long)nx * (y + (long long)ny * z);
  }
  __syncthreads();
  const int wx = (bxhi - bxlo) * b0;  // slab row length
  const int per = rc * b0;             // one block's run
  const int tot = rc * wx;
  const float rwx = 1.f / (float)wx, rper = 1.f / (float)per, rb0 = 1.f / (float)b0;
  const int sx = (int)g.shift[0], xlo = bxlo * b0;
  const long long mrow = (long long)r0 * b0;
  for (long long q = blockIdx.y; q < nvols; q += gridDim.y) {
    const long long vq = q * g.N;
    const long long mq = (c0 + bxlo - start) * nvols * m + q * m + mrow;
    // volume side: element j of the slab is row rr = j / wx, x = xlo + j % wx
    // block side: element j is block bxlo + j / per, run offset j % per
    for (int j0 = threadIdx.x; j0 < tot; j0 += 256 * NB) {
      E v[NB];
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int j = j0 + u * 256;
        if (j < tot) {
          if (!UNLIFT) {
            const int rr = small_div(j, wx, rwx);
            int x = xlo + (j - rr * wx) + sx;
            if (x >= nx) x -= nx;
            v[u] = (NT & 1) ? nt_load(src + vq + rowoff[rr] + x) : src[vq + rowoff[rr] + x];
          } else {
            const int bb = small_div(j, per, rper);
            const E* sp = src + mq + (long long)bb * nvols * m + (j - bb * per);
            v[u] = (NT & 1) ? nt_load(sp) : *sp;
          }
        }
      }
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int j = j0 + u * 256;
        if (j < tot) {
          int t;
          if (!UNLIFT) {
            const int rr = small_div(j, wx, rwx);
            t = rr * WP + (j - rr * wx);
          } else {
            const int bb = small_div(j, per, rper), e = j - bb * per;
            const int rr = small_div(e, b0, rb0);
            t = rr * WP + bb * b0 + (e - rr * b0);
          }
          tile[t] = v[u];
        }
      }
    }
    __syncthreads();
    for (int j0 = threadIdx.x; j0 < tot; j0 += 256 * NB) {
      E v[NB];
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int j = j0 + u * 256;
        if (j < tot) {
          int t;
          if (!UNLIFT) {
            const int bb = small_div(j, per, rper), e = j - bb * per;
            const int rr = small_div(e, b0, rb0);
            t = rr * WP + bb * b0 + (e - rr * b0);
          } else {
            const int rr = small_div(j, wx, rwx);
            t = rr * WP + (j - rr * wx);
          }
          v[u] = tile[t];
        }
      }
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int j = j0 + u * 256;
        if (j < tot) {
          if (!UNLIFT) {
            const int bb = small_div(j, per, rper);
            E* dp = dst + mq + (long long)bb * nvols * m + (j - bb * per);
            if (NT & 2) nt_store(v[u], dp); else *dp = v[u];
          } else {
            const int rr = small_div(j, wx, rwx);
            int x = xlo + (j - rr * wx) + sx;
            if (x >= nx) x -= nx;
            E* dp = dst + vq + rowoff[rr] + x;
            if (NT & 2) nt_store(v[u], dp); else *dp = v[u];
          }
        }
      }
    }
    __syncthreads();
  }
}

// byte-granular fallback for element sizes other than 1/2/4/8/16
__global__ void k_lift_bytes(const unsigned char* __restrict__ vols, unsigned char* __restrict__ mats, GeomArgs g,
                             long long nvols, long long start, long long total, long long eb, int dir) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    long long byte = e % eb, el = e / eb;
    long long i = el % g.m, rest = el / g.m;
    long long q = rest % nvols, c = rest / nvols;
    long long va = (q * g.N + block_voxel_addr(g, start + c, i)) * eb + byte;
    if (dir == 0)
      mats[e] = vols[va];
    else
      ((unsigned char*)vols)[va] = mats[e];
  }
}

// ------------------------------------------------------ subspace projection
template <int K>
__global__ void __launch_bounds__(256) k_project(const float2* __restrict__ X, float2* __restrict__ Cp,
                                                 const float2* __restrict__ S, long long N, int nq) {
  for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < N; v += (long long)gridDim.x * blockDim.x) {
    float2 acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = make_float2(0.f, 0.f);
    for (int q = 0; q < nq; ++q) {
      float2 x = StoreF32::ldnt(X + (long long)q * N + v);
#pragma unroll
      for (int k = 0; k < K; ++k) {
        float2 s = S[q + (long long)k * nq];
        acc[k].x += x.x * s.x - x.y * s.y;
        acc[k].y += x.x * s.y + x.y * s.x;
      }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) Cp[(long long)k * N + v] = acc[k];
  }
}

// X[q][v] = sum_k L[k][v] conj(S[q][k]) for every voxel v inside the shifted
// block grid (the back-projection of dti_back_project + conj_transpose,
// lib.rs:288-289), streamed in voxel order; voxels the grid leaves out keep
// their value (block.rs:112-118 skips the remainder).
// X is stored as ST (complex64 or half-precision pairs); `full` skips the coverage test
// when the grid covers every voxel.
template <int K, class ST>
__global__ void __launch_bounds__(256) k_backproject(const float2* __restrict__ L, typename ST::T* __restrict__ X,
                                                     const float2* __restrict__ S, int nx, int ny, int nz, int cx,
                                                     int cy, int cz, int sx, int sy, int sz, int nq, int full,
                                                     float scale) {
  const unsigned N = (unsigned)nx * ny * nz;
  for (unsigned v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
    if (!full) {
      int x = (int)(v % (unsigned)nx), y = (int)((v / (unsigned)nx) % (unsigned)ny), z = (int)(v / ((unsigned)nx * ny));
      int rx = x - sx, ry = y - sy, rz = z - sz;
      rx += rx < 0 ? nx : 0;
      ry += ry < 0 ? ny : 0;
      rz += rz < 0 ? nz : 0;
      if (rx >= cx || ry >= cy || rz >= cz) continue;
    }
    float2 c[K];
#pragma unroll
    for (int k = 0; k < K; ++k) c[k] = L[(size_t)k * N + v];
    for (int q = 0; q < nq; ++q) {
      float2 acc = make_float2(0.f, 0.f);
#pragma unroll
      for (int k = 0; k < K; ++k) {
        float2 s = S[q + k * nq];
        acc.x += c[k].x * s.x + c[k].y * s.y;
        acc.y += c[k].y * s.x - c[k].x * s.y;
      }
      ST::st(X + (size_t)q * N + v, cscale(acc, scale));
    }
  }
}

// The fused LLR prox of block slot `slot` of A (one workgroup of 256 threads;
// `smem` is the kernel's dynamic LDS).
__device__ void llr_block_body(const BlockArgs& A, long long slot, float2* smem) {
  const int tid = threadIdx.x;
  const int nc = A.nc;
  const int m = (int)A.g.m;  // < 2^20 (geom_in_range)
  const long long b = A.block0 + slot;
  const int mp = m + 1;  // padded column stride in LDS
  float2* Fs = smem;                                      // nc * mp (when staged)
  float2* Gp = smem + (A.stage_lds ? nc * mp : 0);        // 256 partial Gram entries
  float2* Ws = Gp + 256;                                  // 64: W
  float2* Ps = Ws + 64;                                   // nc * nq: W S^H

  const BlockRows br(A.g, b);
  const long long cs = col_stride(A);
  if (A.stage_lds) {
    // two rows per thread at a time, every load issued before the LDS stores
    for (int i0 = tid; i0 < m; i0 += 512) {
      float2 v[2][8];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int i = i0 + 256 * h;
        const long long base = i < m ? row_base(A, br, b, i) : -1;
#pragma unroll
        for (int col = 0; col < 8; ++col)
          v[h][col] = (col < nc && base >= 0) ? A.F[base + col * cs] : make_float2(0.f, 0.f);
      }
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int i = i0 + 256 * h;
        if (i >= m) continue;
#pragma unroll
        for (int col = 0; col < 8; ++col)
          if (col < nc) Fs[col * mp + i] = v[h][col];
      }
    }
    __syncthreads();
  }
  // Gram: entry e = (i <= j) for thread t % NE, rows interleaved by t / NE
  const int NE = nc * (nc + 1) / 2;
  const int NH = 256 / NE;
  {
    int e = tid % NE, h = tid / NE;
    int i = 0, j = 0;
    {
      int rem = e;
      for (int jj = 0; jj < nc; ++jj) {
        if (rem <= jj) {
          i = rem;
          j = jj;
          break;
        }
        rem -= jj + 1;
      }
    }
    float2 acc = make_float2(0.f, 0.f);
    if (h < NH) {
      for (int r = h; r < m; r += NH) {
        float2 fi, fj;
        if (A.stage_lds) {
          fi = Fs[i * mp + r];
          fj = Fs[j * mp + r];
        } else {
          fi = load_elem(A, br, b, r, i);
          fj = load_elem(A, br, b, r, j);
        }
        acc.x += fi.x * fj.x + fi.y * fj.y;
        acc.y += fi.x * fj.y - fi.y * fj.x;
      }
    }
    Gp[tid] = acc;
  }
  __syncthreads();
  // wave 0: reduce partials, Jacobi, weights, W
  if (tid < 64) {
    const int lane = tid;
    const int r = lane & 7, c = lane >> 3;
    float2 a = make_float2(0.f, 0.f);
    if (r < nc && c < nc) {
      int i = min(r, c), j = max(r, c);
      int e = j * (j + 1) / 2 + i;
      float2 s = make_float2(0.f, 0.f);
      for (int h = 0; h < NH; ++h) s = cadd(s, Gp[h * NE + e]);
      a = (r <= c) ? s : cconj(s);
      if (r == c) a.y = 0.f;
    }
    float2 vv;
    wave_jacobi8<float>(a, vv, lane, 16, 2.5e-7f);
    float lam[8];
#pragma unroll
    for (int l = 0; l < 8; ++l) lam[l] = __shfl(a.x, 9 * l, 64);
    float lmax = 0.f;
#pragma unroll
    for (int l = 0; l < 8; ++l)
      if (l < nc) lmax = fmaxf(lmax, lam[l]);
    if (lane == 0) {
      float smax = sqrtf(fmaxf(lmax, 0.f));
      if (A.smax_bits) atomicMax(A.smax_bits, __float_as_uint(smax));
    }
    if (A.svals && lane < 8) A.svals[slot * 8 + lane] = sqrtf(fmaxf(__shfl(a.x, 9 * lane, 64), 0.f));
    if (A.mode != 2) {
      float gl[8];
#pragma unroll
      for (int l = 0; l < 8; ++l) gl[l] = shrink_weight(lam, l, nc, A.mode, A.rank, A.lambda);
      float2 w = make_float2(0.f, 0.f);
#pragma unroll
      for (int l = 0; l < 8; ++l) {
        float2 vrl = shfl2(vv, r + 8 * l);
        float2 vcl = shfl2(vv, c + 8 * l);
        float2 t = cmulc(vrl, vcl);
        w.x += gl[l] * t.x;
        w.y += gl[l] * t.y;
      }
      Ws[lane] = w;  // W[r][c]
    }
  }
  if (A.mode == 2) return;
  __syncthreads();
  if (A.S) {
    // P[k][q] = sum_c W[k][c] conj(S[q][c])
    for (int e = tid; e < nc * A.nq; e += 256) {
      int k = e % nc, q = e / nc;
      float2 acc = make_float2(0.f, 0.f);
      for (int cc = 0; cc < nc; ++cc) acc = cadd(acc, cmulc(Ws[k + 8 * cc], A.S[q + (long long)cc * A.nq]));
      Ps[k + nc * q] = acc;
    }
    __syncthreads();
  }
  // apply: each thread handles rows r = tid, tid+256, ...
  for (int rr = tid; rr < m; rr += 256) {
    const long long base = row_base(A, br, b, rr);
    if (base < 0) continue;  // outside a partial edge block
    // (two loops: a load through a pointer chosen between LDS and global
    // memory is a generic load)
    float2 f[8];
    if (A.stage_lds) {
#pragma unroll
      for (int k = 0; k < 8; ++k) f[k] = k < nc ? Fs[k * mp + rr] : make_float2(0.f, 0.f);
    } else {
#pragma unroll
      for (int k = 0; k < 8; ++k) f[k] = k < nc ? A.F[base + k * cs] : make_float2(0.f, 0.f);
    }
    if (A.S) {
      const long long addr = base;  // layout 0 (a subspace output needs voxel planes)
      for (int q = 0; q < A.nq; ++q) {
        float2 acc = make_float2(0.f, 0.f);
#pragma unroll
        for (int k = 0; k < 8; ++k)
          if (k < nc) acc = cadd(acc, cmul(f[k], Ps[k + nc * q]));
        A.out[(long long)q * A.g.N + addr] = acc;
      }
    } else {
      for (int cc = 0; cc < nc; ++cc) {
        float2 acc = make_float2(0.f, 0.f);
#pragma unroll
        for (int k = 0; k < 8; ++k)
          if (k < nc) acc = cadd(acc, cmul(f[k], Ws[k + 8 * cc]));
        A.out[base + cc * cs] = acc;
      }
    }
  }
}

__global__ void __launch_bounds__(256) k_llr_block(BlockArgs A) {
  extern __shared__ float2 smem[];
  const long long slot = xcd_slot(A.per_xcd);
  if (slot >= A.nblocks) return;
  llr_block_body(A, slot, smem);
}

// One member of a batch of reconstructions for the persistent LLR launch.
// The shift of the current iteration is read from the member's device table at
// the index held by the device iteration counter, so a captured graph replays
// every iteration unchanged.
struct MemberBlocks {
  BlockArgs A;          // g.shift is filled in per iteration
  long long first;      // first batch-wide item of this member
  const int* shifts;    // n_iter x 3
};

// Persistent LLR launch over the blocks of every member: workgroup g owns the
// contiguous item range of chunk xcd_slot(per_xcd), so each XCD sweeps
// neighbouring blocks of one member.
__global__ void __launch_bounds__(256) k_llr_block_multi(const MemberBlocks* __restrict__ M, int nm, long long total,
                                                         long long per_wg, int per_xcd, const unsigned* iter) {
  extern __shared__ float2 smem[];
  const long long chunk = xcd_slot(per_xcd);
  const long long lo = chunk * per_wg, hi = min(total, lo + per_wg);
  const unsigned it = *iter;
  for (long long item = lo; item < hi; ++item) {
    int j = 0;
    while (j + 1 < nm && M[j + 1].first <= item) ++j;
    BlockArgs A = M[j].A;
    for (int a = 0; a < 3; ++a) A.g.shift[a] = M[j].shifts[3 * (size_t)it + a];
    llr_block_body(A, item - M[j].first, smem);
    __syncthreads();  // LDS is reused by the next item
  }
}

}  // namespace

namespace llr {

llr_status lift_device(llr_ctx* ctx, const void* vols, void* mats, size_t nvols, size_t eb, const size_t* vol,
                       const size_t* blk, const size_t* shift, size_t start, size_t n, bool unlift) {
  GeomArgs g = make_geom(vol, blk, shift);
  if (!geom_in_range(g)) return fail(LLR_ERR_UNSUPPORTED, "block geometry beyond the lift kernel's index range");
  long long total = (long long)n * (long long)nvols * g.m;
  if (total == 0) return LLR_OK;
  hipStream_t s = ctx->stream;
#define LIFT_CASE(E)                                                                                          \
  if (unlift)                                                                                                 \
    hipLaunchKernelGGL((k_lift_rows<E, true>), grid, dim3(256), 0, s, (const E*)mats, (E*)(void*)vols, g,     \
                       (long long)nvols, (long long)start, (long long)n);                                    \
  else                                                                                                        \
    hipLaunchKernelGGL((k_lift_rows<E, false>), grid, dim3(256), 0, s, (const E*)vols, (E*)mats, g,           \
                       (long long)nvols, (long long)start, (long long)n);
  // grid.x: the block-grid rows the range touches; grid.y: volume groups of 4;
  // grid.z: runs of block rows (about 16 rows each)
  const long long gx = g.grid[0];
  const long long rows = (long long)((start + n - 1) / gx) - (long long)(start / gx) + 1;
  const long long nr = g.blk[1] * g.blk[2];
  const dim3 grid((unsigned)rows, (unsigned)std::min<size_t>((nvols + 3) / 4, 65535),
                  (unsigned)std::max<long long>(1, std::min<long long>(nr / 16, 64)));
  // Slab shapes: the side a direction writes gets the whole runs. Lift writes
  // the block side and takes 64 block rows of 8 blocks (a whole 4 KB block
  // run per volume for 8^3 complex64 blocks); un_lift writes the volume side
  // and takes 8 block rows of the whole grid row (whole volume rows). Either
  // shape falls back to the other, then to the grid-row kernel, when its LDS
  // tile would exceed 64 KiB.
  auto slab_bytes = [&](int rc, long long xb) {
    return (size_t)rc * (size_t)(std::min<long long>(xb, g.grid[0]) * g.blk[0] + LB_PAD) * eb;
  };
  const bool tall_ok = slab_bytes(64, 8) <= 65536, wide_ok = slab_bytes(8, g.grid[0]) <= 65536;
  int shape = 0;  // 1: 64 block rows x 8 blocks; 2: 8 block rows x the grid row
  if (eb == 4 || eb == 8 || eb == 16) shape = unlift ? (wide_ok ? 2 : tall_ok ? 1 : 0) : (tall_ok ? 1 : wide_ok ? 2 : 0);
  if (shape) {
    const int rc = shape == 1 ? 64 : 8;
    const int xbw = shape == 1 ? (int)std::min<long long>(8, g.grid[0]) : (int)g.grid[0];
    const int nxc = (int)((g.grid[0] + xbw - 1) / xbw);
    const size_t slab_lds = slab_bytes(rc, xbw);
    // volumes per workgroup: 1 by default (256x256x128x67 complex64, 8^3 blocks,
    // swept over 1-67: lift 4.93 / un_lift 5.28 TB/s at 1 against 4.73 / 5.04 at 4)
    const size_t vpw = ctx->lift_vpw > 0 ? (size_t)ctx->lift_vpw : 1;
    const dim3 sgrid((unsigned)(rows * nxc), (unsigned)std::min<size_t>((nvols + vpw - 1) / vpw, 65535),
                     (unsigned)((nr + rc - 1) / rc));
    // non-temporal stores (lift) and loads and stores (un_lift): each element
    // is moved once; 4.73 -> 4.87-5.05 and 5.17 -> 5.33-5.40 TB/s at
    // 256x256x128x67 complex64, 8^3 blocks
#define SLAB_LAUNCH(E, RC)                                                                                         \
  if (unlift)                                                                                                      \
    hipLaunchKernelGGL((k_lift_slab<E, true, RC, 3>), sgrid, dim3(256), slab_lds, s, (const E*)mats,               \
                       (E*)(void*)vols, g, (long long)nvols, (long long)start, (long long)n, nxc, xbw);            \
  else                                                                                                             \
    hipLaunchKernelGGL((k_lift_slab<E, false, RC, 2>), sgrid, dim3(256), slab_lds, s, (const E*)vols, (E*)mats, g, \
                       (long long)nvols, (long long)start, (long long)n, nxc, xbw);
#define SLAB_CASE(E)     \
  if (shape == 1) {      \
    SLAB_LAUNCH(E, 64)   \
  } else {               \
    SLAB_LAUNCH(E, 8)    \
  }
    switch (eb) {
      case 4: SLAB_CASE(uint32_t); break;
      case 8: SLAB_CASE(uint2); break;
      default: SLAB_CASE(uint4); break;
    }
#undef SLAB_CASE
#undef SLAB_LAUNCH
    LLR_HIP(hipGetLastError());
    return LLR_OK;
  }
  switch (eb) {
    case 1: LIFT_CASE(uint8_t); break;
    case 2: LIFT_CASE(uint16_t); break;
    case 4: LIFT_CASE(uint32_t); break;
    case 8: LIFT_CASE(uint2); break;
    case 16: LIFT_CASE(uint4); break;
    default: {
      long long tb = total * (long long)eb;
      hipLaunchKernelGGL(k_lift_bytes, dim3(grid_for((size_t)tb, 256)), dim3(256), 0, s, (const unsigned char*)vols,
                         (unsigned char*)mats, g, (long long)nvols, (long long)start, tb, (long long)eb, unlift ? 1 : 0);
    }
  }
#undef LIFT_CASE
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

llr_status backproject_device(llr_ctx* ctx, const float2* C, void* X, const float2* S, const size_t* vol,
                              const size_t* blk, const size_t* shift, int nq, int K, bool partial, bool h16,
                              float scale) {
  size_t N = vol[0] * vol[1] * vol[2];
  if (N >= 0xFFFFFFFFull) return fail(LLR_ERR_UNSUPPORTED, "volume larger than 2^32 voxels");
  int cov[3], sh[3];
  for (int a = 0; a < 3; ++a) {
    // extent covered by the blocks: whole blocks only, or everything with partial edge blocks
    cov[a] = (int)(partial ? vol[a] : (vol[a] / blk[a]) * blk[a]);
    sh[a] = (int)(shift[a] % vol[a]);
  }
  const int full = cov[0] == (int)vol[0] && cov[1] == (int)vol[1] && cov[2] == (int)vol[2];
  unsigned grid = grid_for(N, 256, 65536u);
#define BP_CASE(KK)                                                                                                 \
  case KK:                                                                                                          \
    if (h16)                                                                                                       \
      hipLaunchKernelGGL((k_backproject<KK, StoreF16>), dim3(grid), dim3(256), 0, ctx->stream, C, (ch16*)X, S,    \
                         (int)vol[0], (int)vol[1], (int)vol[2], cov[0], cov[1], cov[2], sh[0], sh[1], sh[2], nq, full, scale); \
    else                                                                                                            \
      hipLaunchKernelGGL((k_backproject<KK, StoreF32>), dim3(grid), dim3(256), 0, ctx->stream, C, (float2*)X, S,    \
                         (int)vol[0], (int)vol[1], (int)vol[2], cov[0], cov[1], cov[2], sh[0], sh[1], sh[2], nq, full, scale); \
    break;
  switch (K) {
    BP_CASE(1) BP_CASE(2) BP_CASE(3) BP_CASE(4) BP_CASE(5) BP_CASE(6) BP_CASE(7) BP_CASE(8)
    default: return fail(LLR_ERR_UNSUPPORTED, "subspace rank above 8");
  }
#undef BP_CASE
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

llr_status project_device(llr_ctx* ctx, const float2* X, float2* Cp, const float2* S, size_t N, int nq, int K) {
  unsigned grid = grid_for(N, 256, 65536u);
#define PROJ_CASE(KK) \
  case KK: hipLaunchKernelGGL(k_project<KK>, dim3(grid), dim3(256), 0, ctx->stream, X, Cp, S, (long long)N, nq); break;
  switch (K) {
    PROJ_CASE(1) PROJ_CASE(2) PROJ_CASE(3) PROJ_CASE(4) PROJ_CASE(5) PROJ_CASE(6) PROJ_CASE(7) PROJ_CASE(8)
    PROJ_CASE(9) PROJ_CASE(10) PROJ_CASE(11) PROJ_CASE(12) PROJ_CASE(13) PROJ_CASE(14) PROJ_CASE(15) PROJ_CASE(16)
    default: return fail(LLR_ERR_UNSUPPORTED, "subspace rank above 16");
  }
#undef PROJ_CASE
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

// Shared-memory bytes of k_llr_block for a block of m rows and nc columns.
size_t llr_block_smem(long long m, int nc, int nq, bool stage) {
  size_t f = stage ? (size_t)nc * (size_t)(m + 1) : 0;
  return (f + 256 + 64 + (size_t)8 * (size_t)nq) * sizeof(float2);
}

// Fused LLR prox on blocks [block0, block0 + nblocks).
//   layout 0: F = nc planes of N voxels, blocks are shifted blocks (geom g)
//   layout 1: F = contiguous m x nc matrices
llr_status llr_block_run(llr_ctx* ctx, const float2* F, float2* out, const float2* S, int nq, const size_t* vol,
                         const size_t* blk, const size_t* shift, int nc, int layout, long long block0,
                         long long nblocks, int mode, int rank, float lambda, unsigned* smax_bits, float* svals,
                         bool partial) {
  GeomArgs g = make_geom(vol, blk, shift, partial);
  if (nc < 1 || nc > 8) return fail(LLR_ERR_UNSUPPORTED, "fused LLR kernel handles 1..8 columns");
  if (!geom_in_range(g)) return fail(LLR_ERR_UNSUPPORTED, "block geometry beyond the LLR kernel's index range");
  if (S && layout != 0) return fail(LLR_ERR_ARG, "a subspace output needs the voxel-plane layout");
  if (nblocks <= 0) return LLR_OK;
  BlockArgs A{};
  A.F = F;
  A.out = out;
  A.S = S;
  A.g = g;
  A.nc = nc;
  A.nq = S ? nq : 0;
  A.layout = layout;
  A.nblocks = nblocks;
  A.block0 = block0;
  A.per_xcd = (int)((nblocks + 7) / 8);
  A.mode = mode;
  A.rank = rank;
  A.lambda = lambda;
  A.smax_bits = smax_bits;
  A.svals = svals;
  size_t smem = llr_block_smem(g.m, nc, A.nq, true);
  A.stage_lds = (smem <= 64 * 1024 && ctx->llr_stage != 0) ? 1 : 0;
  if (!A.stage_lds) smem = llr_block_smem(g.m, nc, A.nq, false);
  if (smem > 64 * 1024) return fail(LLR_ERR_UNSUPPORTED, "subspace with too many volumes for the LLR kernel");
  dim3 grid((unsigned)(8 * A.per_xcd));
  hipLaunchKernelGGL(k_llr_block, grid, dim3(256), smem, ctx->stream, A);
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

// Persistent LLR launch over the blocks of several reconstructions.
llr_status llr_multi_plan(MultiPlan& plan, const std::vector<MemberSpec>& members) {
  std::vector<MemberBlocks> h(members.size());
  long long total = 0;
  size_t smem = 0;
  for (size_t j = 0; j < members.size(); ++j) {
    const MemberSpec& ms = members[j];
    if (ms.nc < 1 || ms.nc > 8) return fail(LLR_ERR_UNSUPPORTED, "fused LLR kernel handles 1..8 columns");
    size_t zero[3] = {0, 0, 0};
    GeomArgs g = make_geom(ms.vol, ms.blk, zero, ms.partial);
    if (!geom_in_range(g)) return fail(LLR_ERR_UNSUPPORTED, "block geometry beyond the LLR kernel's index range");
    BlockArgs A{};
    A.F = ms.F;
    A.out = ms.F;
    A.g = g;
    A.nc = ms.nc;
    A.layout = 0;
    A.nblocks = g.grid[0] * g.grid[1] * g.grid[2];
    A.mode = ms.mode;
    A.rank = ms.rank;
    A.lambda = ms.lambda;
    size_t sm = llr_block_smem(g.m, ms.nc, 0, true);
    A.stage_lds = (sm <= 64 * 1024 && ms.stage != 0) ? 1 : 0;
    if (!A.stage_lds) sm = llr_block_smem(g.m, ms.nc, 0, false);
    smem = std::max(smem, sm);
    h[j].A = A;
    h[j].first = total;
    h[j].shifts = ms.d_shifts;
    total += A.nblocks;
  }
  if (smem > 64 * 1024) return fail(LLR_ERR_UNSUPPORTED, "block too large for the LLR kernel");
  LLR_HIP(plan.desc.alloc(h.size() * sizeof(MemberBlocks)));
  LLR_HIP(hipMemcpy(plan.desc.p, h.data(), h.size() * sizeof(MemberBlocks), hipMemcpyHostToDevice));
  plan.nm = (int)h.size();
  plan.total = total;
  plan.smem = smem;
  // 256 CUs x 4 workgroups, dealt to the 8 XCDs in contiguous chunks
  long long G = std::min<long long>(1024, (total + 7) / 8 * 8);
  G = std::max<long long>(G, 8);
  plan.grid = (unsigned)G;
  plan.per_xcd = (int)(G / 8);
  plan.per_wg = (total + G - 1) / G;
  return LLR_OK;
}

llr_status llr_multi_launch(hipStream_t s, const MultiPlan& plan, const unsigned* iter) {
  if (plan.total == 0) return LLR_OK;
  hipLaunchKernelGGL(k_llr_block_multi, dim3(plan.grid), dim3(256), plan.smem, s,
                     (const MemberBlocks*)plan.desc.p, plan.nm, plan.total, plan.per_wg, plan.per_xcd, iter);
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

// Low-rank/SVT of contiguous m x n matrices with 8 < n <= 128 (the wide LLR kernels).
llr_status lowrank_generic_device(llr_ctx* ctx, float2* mats, long long m, int n, long long batch, int mode, int rank,
                                  float lambda, unsigned* smax_bits) {
  if (n > 128) return fail(LLR_ERR_UNSUPPORTED, "more than 128 columns");
  size_t vol[3] = {(size_t)m, 1, 1}, zero[3] = {0, 0, 0};
  return llr_block_run_wide(ctx, mats, vol, vol, zero, n, 1, 0, batch, mode, rank, lambda, smax_bits, false, nullptr);
}

}  // namespace llr
