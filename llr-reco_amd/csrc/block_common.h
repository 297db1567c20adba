// Shifted-block geometry, the wavefront Hermitian Jacobi eigen-solver and the
// block-kernel argument block shared by the FP32 and mixed-precision LLR
// kernels (llr.hip, mixed.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>

#include "common.h"

namespace {

// ------------------------------------------------------------ block geometry
struct GeomArgs {
  long long vol[3], blk[3], grid[3], shift[3];
  long long N;  // voxels per volume
  long long m;  // voxels per block
};

// floor(i / d) for 0 <= i < 2^20 and 1 <= d < 2^20, rd = 1.f / d: the float
// quotient is within 0.2 of i / d, so one step either way corrects it.
__device__ __forceinline__ int small_div(int i, int d, float rd) {
  int q = (int)((float)i * rd);
  q += ((q + 1) * d <= i) ? 1 : 0;
  q -= (q * d > i) ? 1 : 0;
  return q;
}

// Rows of one block. Block b's corner is sub2ind(b, grid) * blk + shift with
// periodic wrap (block.rs:8-28, 122-133); row i is the voxel i = ix + bx (iy +
// by iz) inside it. The corner is found once per block; a row then costs two
// small divisions and three conditional wraps (shift < vol, set by make_geom
// and the shift tables). addr(i) is -1 for a row past the end of the volume
// (possible only on the ceil(vol/blk) grid of partial edge blocks).
struct BlockRows {
  int cx, cy, cz, b0, b01, v0, v1, v2, s0, s1, s2;
  float rb0, rb01;
  __device__ __forceinline__ BlockRows(const GeomArgs& g, long long b) {
    const unsigned g0 = (unsigned)g.grid[0], g01 = (unsigned)(g.grid[0] * g.grid[1]);
    const unsigned ub = (unsigned)b, bz = ub / g01, rr = ub - bz * g01, by = rr / g0, bx = rr - by * g0;
    b0 = (int)g.blk[0];
    b01 = (int)(g.blk[0] * g.blk[1]);
    cx = (int)bx * b0;
    cy = (int)by * (int)g.blk[1];
    cz = (int)bz * (int)g.blk[2];
    v0 = (int)g.vol[0];
    v1 = (int)g.vol[1];
    v2 = (int)g.vol[2];
    s0 = (int)g.shift[0];
    s1 = (int)g.shift[1];
    s2 = (int)g.shift[2];
    rb0 = 1.f / (float)b0;
    rb01 = 1.f / (float)b01;
  }
  __device__ __forceinline__ long long addr(int i) const {
    const int iz = small_div(i, b01, rb01), r = i - iz * b01;
    const int iy = small_div(r, b0, rb0), ix = r - iy * b0;
    int x = cx + ix, y = cy + iy, z = cz + iz;
    if (x >= v0 || y >= v1 || z >= v2) return -1;
    x += s0;
    y += s1;
    z += s2;
    if (x >= v0) x -= v0;
    if (y >= v1) y -= v1;
    if (z >= v2) z -= v2;
    return (long long)x + (long long)v0 * ((long long)y + (long long)v1 * z);
  }
};

// Voxel address of row i of block b, or -1 (one-off uses; kernels that touch a
// block's rows many times keep a BlockRows).
__device__ __forceinline__ long long block_voxel_addr(const GeomArgs& g, long long b, long long i) {
  return BlockRows(g, b).addr((int)i);
}

// Host-side limits of the 32-bit row arithmetic above.
inline bool geom_in_range(const GeomArgs& g) {
  long long blocks = g.grid[0] * g.grid[1] * g.grid[2];
  return g.m < (1LL << 20) && blocks < (1LL << 31) && g.vol[0] < (1LL << 30) && g.vol[1] < (1LL << 30) &&
         g.vol[2] < (1LL << 30) && g.grid[0] * g.blk[0] < (1LL << 30) && g.grid[1] * g.blk[1] < (1LL << 30) &&
         g.grid[2] * g.blk[2] < (1LL << 30);
}

// Block grid floor(vol/blk) (block.rs:112-118), or ceil(vol/blk) with
// truncated edge blocks when `partial`.
inline GeomArgs make_geom(const size_t* vol, const size_t* blk, const size_t* shift, bool partial = false) {
  GeomArgs g;
  for (int a = 0; a < 3; ++a) {
    g.vol[a] = (long long)vol[a];
    g.blk[a] = (long long)blk[a];
    g.grid[a] = (long long)(partial ? (vol[a] + blk[a] - 1) / blk[a] : vol[a] / blk[a]);
    g.shift[a] = (long long)(shift[a] % vol[a]);
  }
  g.N = g.vol[0] * g.vol[1] * g.vol[2];
  g.m = g.blk[0] * g.blk[1] * g.blk[2];
  return g;
}

// --------------------------------------------- complex scalars by precision
template <class R>
struct Cx;
template <>
struct Cx<float> {
  using T = float2;
  __device__ static T mk(float x, float y) { return make_float2(x, y); }
};
template <>
struct Cx<double> {
  using T = double2;
  __device__ static T mk(double x, double y) { return make_double2(x, y); }
};

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ double2 cconjmul(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double cnorm(double2 a) { return a.x * a.x + a.y * a.y; }

__device__ __forceinline__ float2 shfl2(float2 v, int src) {
  return make_float2(__shfl(v.x, src, 64), __shfl(v.y, src, 64));
}
__device__ __forceinline__ double2 shfl2(double2 v, int src) {
  return make_double2(__shfl(v.x, src, 64), __shfl(v.y, src, 64));
}

// ----------------------------------------------------- wavefront Jacobi (8x8)
// partner of index i in round k of the circle-method round robin on 8 indices
__device__ __forceinline__ int rr_partner(int i, int k) {
  if (i == 7) return k;
  if (i == k) return 7;
  int p = (2 * k - i) % 7;
  return p < 0 ? p + 7 : p;
}

template <class R>
struct Rot {
  R cs, sn;
  typename Cx<R>::T u;  // unit phase of A[p][q]
};

// The modulus of A[p][q] is formed scaled (hypot-style): for off-diagonal
// entries near the underflow range sqrt(x^2 + y^2) loses its digits and the
// phase u = a_pq/|a_pq| would not be unit-modulus, making the rotation
// non-unitary. Entries below 1e-9 (FP32) or 1e-18 (FP64) of the diagonal pair
// are left alone.
template <class R>
__device__ __forceinline__ Rot<R> rot_params(R app, R aqq, typename Cx<R>::T apq) {
  Rot<R> r;
  const R skip = sizeof(R) == 4 ? R(1e-9) : R(1e-18);
  R s = fmax(fabs(apq.x), fabs(apq.y));
  if (s == R(0) || s <= skip * (fabs(app) + fabs(aqq))) {
    r.cs = R(1);
    r.sn = R(0);
    r.u = Cx<R>::mk(R(1), R(0));
    return r;
  }
  R xs = apq.x / s, ys = apq.y / s;
  R rn = R(1) / sqrt(xs * xs + ys * ys);
  R a = s / rn;
  r.u = Cx<R>::mk(xs * rn, ys * rn);
  R theta = (aqq - app) / (R(2) * a);
  R at = fabs(theta);
  R t = (at > R(1e18)) ? R(0.5) / theta : ((theta >= R(0) ? R(1) : R(-1)) / (at + sqrt(theta * theta + R(1))));
  r.cs = R(1) / sqrt(t * t + R(1));
  r.sn = t * r.cs;
  return r;
}

// FP32 rotation from the hardware reciprocal and reciprocal square root
// (about 1 ulp each) instead of IEEE-exact division and square root: the
// rotation stays unitary to a few ulp (cs = rsq(1 + t^2), sn = t cs; u from
// rsq of a modulus in [1, 2]) and only the angle carries the approximation,
// which the next rotations of the sweep absorb. Same skip rule and scaling as
// the generic form.
template <>
__device__ __forceinline__ Rot<float> rot_params<float>(float app, float aqq, float2 apq) {
  Rot<float> r;
  const float s = fmaxf(fabsf(apq.x), fabsf(apq.y));
  // below 16 FLT_MIN the hardware reciprocal of s overflows (denormal inputs)
  if (s < 16.f * FLT_MIN || s <= 1e-9f * (fabsf(app) + fabsf(aqq))) {
    r.cs = 1.f;
    r.sn = 0.f;
    r.u = make_float2(1.f, 0.f);
    return r;
  }
  const float is = __builtin_amdgcn_rcpf(s);
  const float xs = apq.x * is, ys = apq.y * is;
  const float q = xs * xs + ys * ys;  // in [1, 2]
  const float rn = __builtin_amdgcn_rsqf(q);
  const float a = s * (q * rn);       // |apq|
  r.u = make_float2(xs * rn, ys * rn);
  const float theta = (aqq - app) * (0.5f * __builtin_amdgcn_rcpf(a));
  const float at = fabsf(theta);
  float t;
  if (at > 1e18f) {
    t = 0.5f * __builtin_amdgcn_rcpf(theta);
  } else {
    const float h = theta * theta + 1.f;
    t = copysignf(__builtin_amdgcn_rcpf(at + h * __builtin_amdgcn_rsqf(h)), theta);
  }
  r.cs = __builtin_amdgcn_rsqf(t * t + 1.f);
  r.sn = t * r.cs;
  return r;
}

// Entries of the 2x2 rotation J for the pair containing index i: J[i][i] and
// J[partner][i].
template <class R>
__device__ __forceinline__ void rot_entries(const Rot<R>& Q, bool i_is_p, typename Cx<R>::T* Jii,
                                            typename Cx<R>::T* Jpi) {
  typename Cx<R>::T cu = Cx<R>::mk(Q.u.x, -Q.u.y);
  if (i_is_p) {
    *Jii = Cx<R>::mk(Q.cs, R(0));
    *Jpi = Cx<R>::mk(-Q.sn * cu.x, -Q.sn * cu.y);
  } else {
    *Jii = Cx<R>::mk(Q.cs * cu.x, Q.cs * cu.y);
    *Jpi = Cx<R>::mk(Q.sn, R(0));
  }
}

template <class R>
__device__ __forceinline__ R wave_sum(R v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Jacobi eigen-solve of the Hermitian 8x8 matrix held one element per lane
// (lane = r + 8c holds A[r][c]). `vv` enters as the starting eigenvector
// matrix when `keep_v`, else the identity. On return a holds the (diagonal)
// eigenvalue matrix and vv the eigenvectors (V[r][c], column c = eigenvector
// c). Convergence is checked once per sweep with a butterfly sum, bitwise
// identical in every lane, so the loop exit is uniform.
template <class R>
__device__ void wave_jacobi8(typename Cx<R>::T& a, typename Cx<R>::T& vv, int lane, int max_sweeps, R tol,
                             bool keep_v = false) {
  using C = typename Cx<R>::T;
  const int r = lane & 7, c = lane >> 3;
  if (!keep_v) vv = Cx<R>::mk(r == c ? R(1) : R(0), R(0));
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    R off = wave_sum<R>(r != c ? cnorm(a) : R(0));
    R tot = wave_sum<R>(cnorm(a));
    if (off <= tol * tol * tot || tot == R(0)) break;
    for (int k = 0; k < 7; ++k) {
      int pr = rr_partner(r, k), pc = rr_partner(c, k);
      int p1 = min(r, pr), q1 = max(r, pr);
      int p2 = min(c, pc), q2 = max(c, pc);
      R app1 = __shfl(a.x, p1 * 9, 64), aqq1 = __shfl(a.x, q1 * 9, 64);
      C apq1 = shfl2(a, p1 + 8 * q1);
      R app2 = __shfl(a.x, p2 * 9, 64), aqq2 = __shfl(a.x, q2 * 9, 64);
      C apq2 = shfl2(a, p2 + 8 * q2);
      C a_r_pc = shfl2(a, r + 8 * pc);
      C a_pr_c = shfl2(a, pr + 8 * c);
      C a_pr_pc = shfl2(a, pr + 8 * pc);
      C v_r_pc = shfl2(vv, r + 8 * pc);
      Rot<R> Rr = rot_params<R>(app1, aqq1, apq1);
      Rot<R> Rc = rot_params<R>(app2, aqq2, apq2);
      C Jcc, Jpcc, Jrr, Jprr;
      rot_entries<R>(Rc, c == p2, &Jcc, &Jpcc);
      rot_entries<R>(Rr, r == p1, &Jrr, &Jprr);
      C aj_r = cadd(cmul(a, Jcc), cmul(a_r_pc, Jpcc));
      C aj_pr = cadd(cmul(a_pr_c, Jcc), cmul(a_pr_pc, Jpcc));
      C an = cadd(cconjmul(Jrr, aj_r), cconjmul(Jprr, aj_pr));
      vv = cadd(cmul(vv, Jcc), cmul(v_r_pc, Jpcc));
      if (r != c && c == pr) an = Cx<R>::mk(R(0), R(0));
      if (r == c) an.y = R(0);
      a = an;
    }
  }
}

// g(sigma_l) for eigenvalue index l, given all eigenvalues lam[0..7]: 1 for the
// top `rank` (mode 0, low_rank.rs:10-69) or max(s - lambda, 0)/s (mode 1, SVT)
__device__ __forceinline__ float shrink_weight(const float* lam, int l, int nc, int mode, int rank, float lambda) {
  if (l >= nc) return 0.f;
  if (mode == 0) {
    int pos = 0;
#pragma unroll
    for (int o = 0; o < 8; ++o)
      if (o < nc) pos += (lam[o] > lam[l] || (lam[o] == lam[l] && o < l)) ? 1 : 0;
    return pos < rank ? 1.f : 0.f;
  }
  float s = sqrtf(fmaxf(lam[l], 0.f));
  return (s > lambda && s > 0.f) ? (s - lambda) / s : 0.f;
}

// ---------------------------------------------------------- block kernels
struct BlockArgs {
  const float2* F;    // planes (layout 0) or matrices (layout 1)
  float2* out;        // planes/matrices; with a subspace: nq image planes
  const float2* S;    // nq x nc col-major, or null
  GeomArgs g;
  int nc;             // columns (<= 8)
  int nq;             // output planes with a subspace
  int layout;         // 0: shifted blocks of planes (plane stride g.N); 1: contiguous m x nc matrices
  long long nblocks;  // blocks/matrices handled by this launch
  long long block0;   // first block index
  int per_xcd;        // blocks per XCD slot (XCD-aware remap)
  int mode;           // 0 rank, 1 SVT, 2 singular values only
  int rank;
  float lambda;
  unsigned* smax_bits;  // atomicMax of the largest singular value (float bits)
  float* svals;         // optional [nblocks][8] singular values (unsorted)
  int stage_lds;        // 1: F block staged in LDS
  unsigned* fp64_count; // mixed form: blocks whose eigen-solve fell back to FP64 (optional)
  float2* prev;         // mixed form: momentum on load, F <- (1 + mom) F - mom prev, prev <- F (null: none)
  float mom;
  int mom_first;        // 1: only store prev (the first step after a set-up)
};

// element address of row i, column col of block b (rows `br` of b when
// layout 0); -1 for a row outside the volume
__device__ __forceinline__ long long elem_addr(const BlockArgs& A, const BlockRows& br, long long b, long long i,
                                               int col) {
  if (A.layout == 1) return b * A.g.m * A.nc + (long long)col * A.g.m + i;
  long long v = br.addr((int)i);
  return v < 0 ? -1 : (long long)col * A.g.N + v;
}
// element of column 0 of row i (a voxel address for layout 0, -1 outside the
// volume); column col is at row_base + col * col_stride
__device__ __forceinline__ long long row_base(const BlockArgs& A, const BlockRows& br, long long b, int i) {
  return A.layout == 1 ? b * A.g.m * A.nc + i : br.addr(i);
}
__device__ __forceinline__ long long col_stride(const BlockArgs& A) { return A.layout == 1 ? A.g.m : A.g.N; }
__device__ __forceinline__ float2 load_elem(const BlockArgs& A, const BlockRows& br, long long b, long long i,
                                            int col) {
  long long e = elem_addr(A, br, b, i, col);
  return e < 0 ? make_float2(0.f, 0.f) : A.F[e];
}

// XCD-aware remap: workgroups are dealt round-robin to the 8 XCDs, so give each
// XCD a contiguous run of blocks (neighbouring blocks share L2 lines).
__device__ __forceinline__ long long xcd_slot(int per_xcd) {
  return (long long)(blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
}

}  // namespace
