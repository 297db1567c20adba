// LLR proximal step for wide blocks: more than 8 matrix columns, up to 128.
// This is the reference's default call without a subspace (lib.rs:290-292,
// the n = nq volumes of every block; low_rank.rs:10-69 through cuSOLVER) and
// the host low_rank.rs drop-in for wide matrices. Three kernels, one
// workgroup of 256 threads per block each:
//   k_gram_wide  G = F^H F. Row chunks of the block are staged in LDS; each
//                thread owns 4x4 tiles of the upper triangle in registers.
//   k_eig_wide   Hermitian cyclic Jacobi on G in LDS, all pairs of a
//                round-robin round rotated at once (columns, then rows),
//                then W = V g(sigma) V^H (or only the largest singular value).
//   k_apply_wide F <- F W in place: W and a row chunk in LDS.
// The Gram route is the same mathematics as the narrow kernel (and as
// gesvda's); the rotations use the same parameters (rot_params, rot_entries).
#include <hip/hip_runtime.h>

#include <cfloat>

#include "block_common.h"
#include "internal.h"

namespace {

constexpr int WMAX = 128;  // columns
constexpr int WLDS = 96;   // columns up to which V shares LDS with the Gram
constexpr int GR = 32;     // rows per staged chunk
constexpr int CP = WMAX + 1;  // chunk row pitch (odd: rows of a column fall on spread banks)

struct WideArgs {
  BlockArgs A;  // geometry, layout, F (read and written in place), nc, mode, rank, lambda, smax_bits
  float2* G;    // [nblocks][nc nc] column-major: the Gram, then W
  float2* Vg;   // [nblocks][nc nc]: V when it does not fit LDS (nc > WLDS), else null
};

// stage rows r0 .. r0 + GR - 1 of block b into ch[row][col] (zeros outside
// the block or the volume)
__device__ __forceinline__ void stage_rows(const BlockArgs& A, const BlockRows& br, long long b, int r0, int m, int n,
                                           float2 (*ch)[CP]) {
  const long long cs = col_stride(A);
  for (int e = threadIdx.x; e < GR * n; e += 256) {
    const int rr = e % GR, c = e / GR, r = r0 + rr;
    float2 v = make_float2(0.f, 0.f);
    if (r < m) {
      const long long base = row_base(A, br, b, r);
      if (base >= 0) v = A.F[base + c * cs];
    }
    ch[rr][c] = v;
  }
}

__global__ void __launch_bounds__(256) k_gram_wide(WideArgs w) {
  const BlockArgs& A = w.A;
  __shared__ float2 ch[GR][CP];
  const int n = A.nc, m = (int)A.g.m;
  const long long slot = blockIdx.x, b = A.block0 + slot;
  const BlockRows br(A.g, b);
  const int nt4 = (n + 3) / 4;
  const int ntiles = nt4 * (nt4 + 1) / 2;
  constexpr int TMAX = ((WMAX / 4) * (WMAX / 4 + 1) / 2 + 255) / 256;  // tiles per thread
  int ti[TMAX], tj[TMAX];
  float2 acc[TMAX][4][4];
#pragma unroll
  for (int k = 0; k < TMAX; ++k) {
    // upper-triangle tile tt = J (J + 1) / 2 + I, I <= J
    int tt = threadIdx.x + 256 * k, J = 0;
    while ((J + 1) * (J + 2) / 2 <= tt) ++J;
    ti[k] = tt < ntiles ? tt - J * (J + 1) / 2 : -1;
    tj[k] = J;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[k][i][j] = make_float2(0.f, 0.f);
  }
  for (int r0 = 0; r0 < m; r0 += GR) {
    __syncthreads();  // the previous chunk is consumed
    stage_rows(A, br, b, r0, m, n, ch);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < TMAX; ++k) {
      if (ti[k] < 0) continue;
      const int c0 = 4 * ti[k], c1 = 4 * tj[k];
      for (int rr = 0; rr < GR; ++rr) {
        float2 a[4], bb[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          a[i] = ch[rr][c0 + i];
          bb[i] = ch[rr][c1 + i];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) {  // conj(a_i) b_j
            acc[k][i][j].x = fmaf(a[i].x, bb[j].x, fmaf(a[i].y, bb[j].y, acc[k][i][j].x));
            acc[k][i][j].y = fmaf(a[i].x, bb[j].y, fmaf(-a[i].y, bb[j].x, acc[k][i][j].y));
          }
      }
    }
  }
  float2* G = w.G + slot * (long long)n * n;
#pragma unroll
  for (int k = 0; k < TMAX; ++k) {
    if (ti[k] < 0) continue;
    const int c0 = 4 * ti[k], c1 = 4 * tj[k];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int gi = c0 + i, gj = c1 + j;
        if (gi >= n || gj >= n) continue;
        float2 v = acc[k][i][j];
        if (gi == gj) v.y = 0.f;
        G[gi + (long long)n * gj] = v;
        if (ti[k] != tj[k]) G[gj + (long long)n * gi] = cconj(v);
      }
  }
}

// round-robin round `rd` of np players (np even): pair k of the round
__device__ __forceinline__ void rr_pair(int np, int rd, int k, int* p, int* q) {
  auto player = [&](int pos) { return pos == 0 ? 0 : 1 + (pos - 1 + rd) % (np - 1); };
  const int a = player(k), b = player(np - 1 - k);
  *p = min(a, b);
  *q = max(a, b);
}

__device__ __forceinline__ float block_sum(float v, float* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

// VG: V in the global work buffer (nc > WLDS)
template <bool VG>
__global__ void __launch_bounds__(256) k_eig_wide(WideArgs w) {
  const BlockArgs& A = w.A;
  extern __shared__ float2 wlds[];
  __shared__ Rot<float> rot[WMAX / 2];
  __shared__ int rp[WMAX / 2], rq[WMAX / 2];
  __shared__ float red[4];
  __shared__ float gl[WMAX];
  __shared__ int act[WMAX];
  __shared__ int nact;
  const int n = A.nc, np = (n + 1) & ~1, ld = np + 1, tid = threadIdx.x;
  const long long slot = blockIdx.x;
  float2* Gg = w.G + slot * (long long)n * n;
  float2* M = wlds;                                                  // np x np, pitch ld
  float2* V = VG ? w.Vg + slot * (long long)np * ld : wlds + np * ld;  // np x np, pitch ld
  for (int e = tid; e < np * np; e += 256) {
    const int i = e % np, j = e / np;
    M[i + ld * j] = (i < n && j < n) ? Gg[i + (long long)n * j] : make_float2(0.f, 0.f);
    V[i + ld * j] = make_float2(i == j ? 1.f : 0.f, 0.f);
  }
  __syncthreads();
  const float tol = 2.5e-7f;
  for (int sweep = 0; sweep < 24; ++sweep) {
    float off = 0.f, tot = 0.f;
    for (int e = tid; e < np * np; e += 256) {
      const int i = e % np, j = e / np;
      const float2 v = M[i + ld * j];
      const float q = v.x * v.x + v.y * v.y;
      tot += q;
      if (i != j) off += q;
    }
    off = block_sum(off, red);
    tot = block_sum(tot, red);
    if (off <= tol * tol * tot || tot == 0.f) break;
    for (int rd = 0; rd < np - 1; ++rd) {
      if (tid < np / 2) {
        int p, q;
        rr_pair(np, rd, tid, &p, &q);
        rp[tid] = p;
        rq[tid] = q;
        rot[tid] = rot_params<float>(M[p + ld * p].x, M[q + ld * q].x, M[p + ld * q]);
      }
      __syncthreads();
      // columns: M <- M J, V <- V J
      for (int e = tid; e < (np / 2) * np; e += 256) {
        const int k = e / np, r = e % np, p = rp[k], q = rq[k];
        float2 Jpp, Jqp, Jqq, Jpq;
        rot_entries<float>(rot[k], true, &Jpp, &Jqp);
        rot_entries<float>(rot[k], false, &Jqq, &Jpq);
        const float2 mp = M[r + ld * p], mq = M[r + ld * q];
        M[r + ld * p] = cadd(cmul(mp, Jpp), cmul(mq, Jqp));
        M[r + ld * q] = cadd(cmul(mq, Jqq), cmul(mp, Jpq));
        const float2 vp = V[r + ld * p], vq = V[r + ld * q];
        V[r + ld * p] = cadd(cmul(vp, Jpp), cmul(vq, Jqp));
        V[r + ld * q] = cadd(cmul(vq, Jqq), cmul(vp, Jpq));
      }
      __syncthreads();
      // rows: M <- J^H M
      for (int e = tid; e < (np / 2) * np; e += 256) {
        const int k = e / np, c = e % np, p = rp[k], q = rq[k];
        float2 Jpp, Jqp, Jqq, Jpq;
        rot_entries<float>(rot[k], true, &Jpp, &Jqp);
        rot_entries<float>(rot[k], false, &Jqq, &Jpq);
        const float2 mp = M[p + ld * c], mq = M[q + ld * c];
        float2 np_ = cadd(cconjmul(Jpp, mp), cconjmul(Jqp, mq));
        float2 nq_ = cadd(cconjmul(Jqq, mq), cconjmul(Jpq, mp));
        if (c == q) np_ = make_float2(0.f, 0.f);
        if (c == p) {
          nq_ = make_float2(0.f, 0.f);
          np_.y = 0.f;
        }
        if (c == q) nq_.y = 0.f;
        M[p + ld * c] = np_;
        M[q + ld * c] = nq_;
      }
      __syncthreads();
    }
  }
  // eigenvalues on the diagonal: the largest singular value, the weights
  if (tid == 0) {
    float lmax = 0.f;
    for (int l = 0; l < n; ++l) lmax = fmaxf(lmax, M[l + ld * l].x);
    if (A.smax_bits) atomicMax(A.smax_bits, __float_as_uint(sqrtf(fmaxf(lmax, 0.f))));
  }
  if (A.mode == 2) return;
  for (int l = tid; l < n; l += 256) {
    const float lam = M[l + ld * l].x;
    float g;
    if (A.mode == 0) {
      int pos = 0;
      for (int o = 0; o < n; ++o) {
        const float lo = M[o + ld * o].x;
        pos += (lo > lam || (lo == lam && o < l)) ? 1 : 0;
      }
      g = pos < A.rank ? 1.f : 0.f;
    } else {
      const float s = sqrtf(fmaxf(lam, 0.f));
      g = (s > A.lambda && s > 0.f) ? (s - A.lambda) / s : 0.f;
    }
    gl[l] = g;
  }
  __syncthreads();
  if (tid == 0) {
    int c = 0;
    for (int l = 0; l < n; ++l)
      if (gl[l] != 0.f) act[c++] = l;
    nact = c;
  }
  __syncthreads();
  // W = V diag(g) V^H over the columns with g != 0
  for (int e = tid; e < n * n; e += 256) {
    const int i = e % n, j = e / n;
    float2 acc = make_float2(0.f, 0.f);
    for (int a = 0; a < nact; ++a) {
      const int l = act[a];
      const float2 t = cmulc(V[i + ld * l], V[j + ld * l]);
      acc.x = fmaf(gl[l], t.x, acc.x);
      acc.y = fmaf(gl[l], t.y, acc.y);
    }
    Gg[i + (long long)n * j] = acc;
  }
}

// F <- F W for the rows of one block, in place (WL: W staged in LDS, else
// read from the work buffer through the caches)
template <bool WL>
__global__ void __launch_bounds__(256) k_apply_wide(WideArgs w) {
  const BlockArgs& A = w.A;
  extern __shared__ float2 alds[];
  const int n = A.nc, m = (int)A.g.m, tid = threadIdx.x;
  const long long slot = blockIdx.x, b = A.block0 + slot;
  const BlockRows br(A.g, b);
  const long long cs = col_stride(A);
  const float2* Wg = w.G + slot * (long long)n * n;
  float2 (*ch)[CP] = reinterpret_cast<float2 (*)[CP]>(alds + (WL ? n * n : 0));  // GR x CP
  if (WL)
    for (int e = tid; e < n * n; e += 256) alds[e] = Wg[e];
  const float2* W = WL ? alds : Wg;  // n x n, column-major
  const int ng = (n + 3) / 4;  // output column groups of 4
  for (int r0 = 0; r0 < m; r0 += GR) {
    __syncthreads();  // W staged / the previous chunk written out
    stage_rows(A, br, b, r0, m, n, ch);
    __syncthreads();
    for (int e = tid; e < GR * ng; e += 256) {
      const int rr = e % GR, jg = e / GR, r = r0 + rr;
      if (r >= m) continue;
      const long long base = row_base(A, br, b, r);
      if (base < 0) continue;  // outside a partial edge block
      const int j0 = 4 * jg;
      float2 acc[4] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
      for (int i = 0; i < n; ++i) {
        const float2 f = ch[rr][i];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const float2 wv = (j0 + j < n) ? W[i + n * (j0 + j)] : make_float2(0.f, 0.f);
          acc[j].x = fmaf(f.x, wv.x, fmaf(-f.y, wv.y, acc[j].x));
          acc[j].y = fmaf(f.x, wv.y, fmaf(f.y, wv.x, acc[j].y));
        }
      }
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (j0 + j < n) A.out[base + (j0 + j) * cs] = acc[j];
    }
  }
}

}  // namespace

namespace llr {

size_t llr_wide_work_elems(int nc, long long nblocks) {
  const size_t np = (size_t)((nc + 1) & ~1);
  size_t e = (size_t)nc * (size_t)nc * (size_t)nblocks;
  if (nc > WLDS) e += np * (np + 1) * (size_t)nblocks;
  return e;
}

// Wide LLR prox on blocks [block0, block0 + nblocks) of F (in place):
//   layout 0: nc planes of vol voxels, shifted blocks; layout 1: contiguous m x nc matrices.
// work holds llr_wide_work_elems(nc, nblocks) complex values (allocated here when null).
llr_status llr_block_run_wide(llr_ctx* ctx, float2* F, const size_t* vol, const size_t* blk, const size_t* shift,
                              int nc, int layout, long long block0, long long nblocks, int mode, int rank,
                              float lambda, unsigned* smax_bits, bool partial, float2* work) {
  if (nc < 1 || nc > WMAX) return fail(LLR_ERR_UNSUPPORTED, "wide LLR kernels handle up to 128 columns");
  GeomArgs g = make_geom(vol, blk, shift, partial);
  if (!geom_in_range(g)) return fail(LLR_ERR_UNSUPPORTED, "block geometry beyond the LLR kernel's index range");
  if (nblocks <= 0) return LLR_OK;
  if (nblocks > 0x7fffffffLL) return fail(LLR_ERR_UNSUPPORTED, "too many blocks for one launch");
  DevBuf own;
  if (!work) {
    LLR_HIP(own.alloc(llr_wide_work_elems(nc, nblocks) * sizeof(float2)));
    work = own.as<float2>();
  }
  WideArgs w{};
  w.A.F = F;
  w.A.out = F;
  w.A.g = g;
  w.A.nc = nc;
  w.A.layout = layout;
  w.A.nblocks = nblocks;
  w.A.block0 = block0;
  w.A.mode = mode;
  w.A.rank = rank;
  w.A.lambda = lambda;
  w.A.smax_bits = smax_bits;
  w.G = work;
  w.Vg = nc > WLDS ? work + (size_t)nc * nc * (size_t)nblocks : nullptr;
  hipStream_t s = ctx->stream;
  const dim3 grid((unsigned)nblocks);
  hipLaunchKernelGGL(k_gram_wide, grid, dim3(256), 0, s, w);
  LLR_HIP(hipGetLastError());
  const size_t np = (size_t)((nc + 1) & ~1);
  if (nc > WLDS) {
    const size_t sm = np * (np + 1) * sizeof(float2);
    LLR_HIP(hipFuncSetAttribute((const void*)k_eig_wide<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    hipLaunchKernelGGL(k_eig_wide<true>, grid, dim3(256), sm, s, w);
  } else {
    const size_t sm = 2 * np * (np + 1) * sizeof(float2);
    if (sm > 64 * 1024)
      LLR_HIP(hipFuncSetAttribute((const void*)k_eig_wide<false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    hipLaunchKernelGGL(k_eig_wide<false>, grid, dim3(256), sm, s, w);
  }
  LLR_HIP(hipGetLastError());
  if (mode != 2) {
    const bool wl = nc <= WLDS;
    const size_t sa = ((wl ? (size_t)nc * nc : 0) + (size_t)GR * CP) * sizeof(float2);
    if (wl) {
      if (sa > 64 * 1024)
        LLR_HIP(hipFuncSetAttribute((const void*)k_apply_wide<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sa));
      hipLaunchKernelGGL(k_apply_wide<true>, grid, dim3(256), sa, s, w);
    } else {
      hipLaunchKernelGGL(k_apply_wide<false>, grid, dim3(256), sa, s, w);
    }
    LLR_HIP(hipGetLastError());
  }
  if (own.p) LLR_HIP(hipStreamSynchronize(s));  // the work buffer is freed on return
  return LLR_OK;
}

}  // namespace llr
