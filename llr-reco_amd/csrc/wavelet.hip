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
// Each level is two passes over its data; a workgroup owns one x-y column of
// tiles and marches along z, keeping the planes consecutive steps share in LDS:
//   k_wav_an3: input planes (with their 2-sample x-y halo) are filtered along x
//     and y into a ring of planes, the z filter turns the ring into all eight
//     subbands of the step's coefficient planes; the seven detail bands (and
//     the approximation at the deepest level) are shrunk and everything is
//     written to the level's band box B (2px x 2ncy x 2ncz; px = ncx rounded
//     up to 16 so band rows start on 128-byte lines; detail halves at px,
//     ncy, ncz).
//   k_wav_syn3: the eight bands' coefficient planes are kept in a ring and
//     synthesised along z, y and x into the step's output planes.
// The input of level l+1 is the approximation box (ncx x ncy x ncz) inside
// level l's band box, read and written in place through its pitches: level
// l+1's synthesis leaves the reconstructed approximation exactly where level
// l's synthesis reads it, so no other buffer is needed. Levels run fine to
// coarse for the analysis and coarse to fine for the synthesis.
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

__device__ __forceinline__ float2 shrink(float2 x, float alpha) {
  float m = sqrtf(x.x * x.x + x.y * x.y);
  if (m == 0.f) return make_float2(0.f, 0.f);
  float s = fmaxf(m - alpha, 0.f) / m;
  return make_float2(x.x * s, x.y * s);
}

__device__ __forceinline__ void fma2(float2& acc, float w, float2 v) {
  acc.x += w * v.x;
  acc.y += w * v.y;
}

// coefficient tile of the analysis (per band) and sample tile of the synthesis
constexpr int TCX = 16, TCY = 8;          // coefficients along x, y
constexpr int IW = 2 * TCX + 2, IH = 2 * TCY + 2;
constexpr int TX = 32, TY = 16;           // samples along x, y
constexpr int NX = TX / 2 + 1, NY = TY / 2 + 1;

// A box of complex samples or coefficients: element (x, y, z) of volume v at
// base + v*vp + z*p2 + y*p1 + x.
struct Box {
  float2* p;
  long long p1, p2, vp;
};

struct AnArgs {
  Box in;
  int ax, ay, az;            // input extent
  Box B;                     // band box of this level
  int px, ncx, ncy, ncz;     // detail offsets (x: px) and coefficients per band
  int tiles_x, tiles_y;
  int deepest;
  float alpha;
  Db2F f;
};

// Analysis, marching along z. A workgroup owns one (TCX x TCY) column of
// coefficient tiles of one volume and steps through z, TCZ coefficient planes
// per step. Each input plane is loaded once, filtered along x and y once, and
// kept in a ring of 2*TCZ + 2 y-filtered planes (plane L in slot (L + 2) mod
// ring), from which the z filter reads the 4 planes each coefficient plane
// needs; the next step's 2*TCZ input planes are loaded into registers while
// the current step filters along z and stores.
constexpr int AN_TCZ = 2;
constexpr int AN_NEW = 2 * AN_TCZ;             // input planes loaded per step
constexpr int AN_RING = 2 * AN_TCZ + 2;        // y-filtered planes kept
constexpr int AN_PLANE = IH * IW;              // staged input plane
constexpr int AN_NLD = (AN_NEW * AN_PLANE + 255) / 256;
constexpr int AN_SMEM = AN_NEW * AN_PLANE + AN_NEW * IH * 2 * TCX + AN_RING * 2 * TCY * 2 * TCX;

// grid.x = tiles_x * tiles_y, grid.y = volumes
__global__ void __launch_bounds__(256) k_wav_an3(AnArgs a) {
  extern __shared__ float2 sm[];
  float2* s_in = sm;                                 // [AN_NEW][IH][IW]
  float2* s_x = s_in + AN_NEW * AN_PLANE;            // [AN_NEW][IH][2 TCX]: x-lo | x-hi
  float2* s_y = s_x + AN_NEW * IH * 2 * TCX;         // ring [AN_RING][2 TCY][2 TCX]: y-lo | y-hi
  const int tid = threadIdx.x;
  const int tx = blockIdx.x % a.tiles_x, ty = blockIdx.x / a.tiles_x;
  const int cx0 = tx * TCX, cy0 = ty * TCY;
  const long long v = blockIdx.y;
  const float2* src = a.in.p + v * a.in.vp;
  float2* dst = a.B.p + v * a.B.vp;
  float2 ld[AN_NLD];
  // planes L0 .. L0 + n - 1 (logical z, reflected on load) into registers
  auto load = [&](int L0, int n) {
#pragma unroll
    for (int q = 0; q < AN_NLD; ++q) {
      const int e = tid + 256 * q;
      if (e < n * AN_PLANE) {
        const int i = e % IW, r = e / IW;
        const int j = r % IH, k = r / IH;
        const int gx = sym(2 * cx0 - 2 + i, a.ax), gy = sym(2 * cy0 - 2 + j, a.ay), gz = sym(L0 + k, a.az);
        ld[q] = src[(long long)gz * a.in.p2 + (long long)gy * a.in.p1 + gx];
      }
    }
  };
  // registers -> LDS, then x and y filters of the n planes into ring slots of L0 ..
  auto filter_xy = [&](int L0, int n) {
#pragma unroll
    for (int q = 0; q < AN_NLD; ++q) {
      const int e = tid + 256 * q;
      if (e < n * AN_PLANE) s_in[e] = ld[q];
    }
    __syncthreads();
    for (int e = tid; e < n * IH * TCX; e += 256) {
      const int c = e % TCX, row = e / TCX;
      float2 sa = make_float2(0.f, 0.f), sd = make_float2(0.f, 0.f);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const float2 x = s_in[row * IW + 2 * c + 3 - k];
        fma2(sa, a.f.dlo[k], x);
        fma2(sd, a.f.dhi[k], x);
      }
      s_x[row * 2 * TCX + c] = sa;
      s_x[row * 2 * TCX + TCX + c] = sd;
    }
    __syncthreads();
    for (int e = tid; e < n * TCY * 2 * TCX; e += 256) {
      const int c2 = e % (2 * TCX), r = e / (2 * TCX);
      const int cy = r % TCY, k = r / TCY;
      float2 sa = make_float2(0.f, 0.f), sd = make_float2(0.f, 0.f);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float2 x = s_x[(k * IH + 2 * cy + 3 - q) * 2 * TCX + c2];
        fma2(sa, a.f.dlo[q], x);
        fma2(sd, a.f.dhi[q], x);
      }
      const int slot = (L0 + k + 2) % AN_RING;
      s_y[(slot * 2 * TCY + cy) * 2 * TCX + c2] = sa;
      s_y[(slot * 2 * TCY + TCY + cy) * 2 * TCX + c2] = sd;
    }
    __syncthreads();
  };
  // the two planes below the first step (the z halo)
  load(-2, 2);
  filter_xy(-2, 2);
  const int steps = (a.ncz + AN_TCZ - 1) / AN_TCZ;
  load(0, AN_NEW);
  for (int st = 0; st < steps; ++st) {
    const int cz0 = st * AN_TCZ;
    filter_xy(2 * cz0, AN_NEW);
    if (st + 1 < steps) load(2 * cz0 + AN_NEW, AN_NEW);
    // z: coefficient plane cz reads planes 2cz-2 .. 2cz+1 (slots 2cz .. 2cz+3 mod ring)
    for (int e = tid; e < AN_TCZ * 2 * TCY * 2 * TCX; e += 256) {
      const int c2 = e % (2 * TCX), r = e / (2 * TCX);
      const int r2 = r % (2 * TCY), cz = r / (2 * TCY);
      const int czz = cz0 + cz;
      float2 sa = make_float2(0.f, 0.f), sd = make_float2(0.f, 0.f);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int slot = (2 * czz + 3 - q) % AN_RING;
        const float2 x = s_y[(slot * 2 * TCY + r2) * 2 * TCX + c2];
        fma2(sa, a.f.dlo[q], x);
        fma2(sd, a.f.dhi[q], x);
      }
      const bool xh = c2 >= TCX, yh = r2 >= TCY;
      const int cx = cx0 + (xh ? c2 - TCX : c2), cy = cy0 + (yh ? r2 - TCY : r2);
      if (cx >= a.ncx || cy >= a.ncy || czz >= a.ncz) continue;
      if (xh || yh || a.deepest) sa = shrink(sa, a.alpha);
      sd = shrink(sd, a.alpha);
      const long long o = (long long)(yh ? a.ncy + cy : cy) * a.B.p1 + (xh ? a.px + cx : cx);
      dst[(long long)czz * a.B.p2 + o] = sa;
      dst[(long long)(a.ncz + czz) * a.B.p2 + o] = sd;
    }
    // the ring slots the next step overwrites were read above
    __syncthreads();
  }
}

struct SynArgs {
  Box B;                     // band box of this level
  int px, ncx, ncy, ncz;
  Box out;                   // samples of this level's input box
  int ax, ay, az;
  int tiles_x, tiles_y;
  Db2F f;
};

// Synthesis, marching along z: a workgroup owns one (TX x TY) column of sample
// tiles of one volume and steps through z, SYN_TZ samples per step. The
// coefficient planes of the eight bands are kept in a ring of SYN_NZ planes
// (plane c in slot c mod ring); a step needs planes z0/2 .. z0/2 + SYN_TZ/2,
// of which SYN_TZ/2 are new, loaded into registers during the previous step.
constexpr int SYN_TZ = 4;
constexpr int SYN_NZ = SYN_TZ / 2 + 1;
constexpr int SYN_NEW = SYN_TZ / 2;
constexpr int SYN_PLANE = NY * NX;                         // one band's coefficient plane of a tile
constexpr int SYN_NLD = (8 * SYN_NEW * SYN_PLANE + 255) / 256;
constexpr int SYN_SMEM = 8 * SYN_NZ * SYN_PLANE + 4 * SYN_TZ * SYN_PLANE + 2 * SYN_TZ * TY * NX;

// grid.x = tiles_x * tiles_y, grid.y = volumes
__global__ void __launch_bounds__(256) k_wav_syn3(SynArgs a) {
  extern __shared__ float2 sm[];
  float2* s_b = sm;                                  // ring [band][slot][iy][ix], band = xb + 2 yb + 4 zb
  float2* s_z = s_b + 8 * SYN_NZ * SYN_PLANE;        // [xb + 2 yb][sz][iy][ix]
  float2* s_y = s_z + 4 * SYN_TZ * SYN_PLANE;        // [xb][sz][sy][ix]
  const int tid = threadIdx.x;
  const int tx = blockIdx.x % a.tiles_x, ty = blockIdx.x / a.tiles_x;
  const int x0 = tx * TX, y0 = ty * TY;
  const int cx0 = x0 / 2, cy0 = y0 / 2;
  const long long v = blockIdx.y;
  const float2* B = a.B.p + v * a.B.vp;
  float2 ld[SYN_NLD];
  // coefficient planes c0 .. c0 + n - 1 of all eight bands into registers
  auto load = [&](int c0, int n) {
#pragma unroll
    for (int q = 0; q < SYN_NLD; ++q) {
      const int e = tid + 256 * q;
      ld[q] = make_float2(0.f, 0.f);
      if (e < 8 * n * SYN_PLANE) {
        const int ix = e % NX, r = e / NX;
        const int iy = r % NY, r2 = r / NY;
        const int k = r2 % n, band = r2 / n;
        const int cx = cx0 + ix, cy = cy0 + iy, cz = c0 + k;
        if (cx < a.ncx && cy < a.ncy && cz < a.ncz) {
          const long long gx = (band & 1) ? a.px + cx : cx;
          const long long gy = (band & 2) ? a.ncy + cy : cy;
          const long long gz = (band & 4) ? a.ncz + cz : cz;
          ld[q] = B[gz * a.B.p2 + gy * a.B.p1 + gx];
        }
      }
    }
  };
  auto store = [&](int c0, int n) {
#pragma unroll
    for (int q = 0; q < SYN_NLD; ++q) {
      const int e = tid + 256 * q;
      if (e < 8 * n * SYN_PLANE) {
        const int ixy = e % SYN_PLANE, r2 = e / SYN_PLANE;
        const int k = r2 % n, band = r2 / n;
        s_b[(band * SYN_NZ + (c0 + k) % SYN_NZ) * SYN_PLANE + ixy] = ld[q];
      }
    }
  };
  float2* dst = a.out.p + v * a.out.vp;
  const int steps = (a.az + SYN_TZ - 1) / SYN_TZ;
  load(0, 1);
  store(0, 1);
  load(1, SYN_NEW);
  for (int st = 0; st < steps; ++st) {
    const int z0 = st * SYN_TZ, cz0 = z0 / 2;
    store(cz0 + 1, SYN_NEW);
    __syncthreads();
    if (st + 1 < steps) load(cz0 + 1 + SYN_NEW, SYN_NEW);
    // along z: sample z0+sz from coefficient planes idx/2, idx = z0+sz+2-j even, inside [0, 2ncz)
    for (int e = tid; e < 4 * SYN_TZ * SYN_PLANE; e += 256) {
      const int ixy = e % SYN_PLANE, r = e / SYN_PLANE;
      const int sz = r % SYN_TZ, pair = r / SYN_TZ;
      float2 acc = make_float2(0.f, 0.f);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int idx = z0 + sz + 2 - j;
        if (idx >= 0 && idx < 2 * a.ncz && (idx & 1) == 0) {
          const int slot = (idx >> 1) % SYN_NZ;
          fma2(acc, a.f.rlo[j], s_b[(pair * SYN_NZ + slot) * SYN_PLANE + ixy]);
          fma2(acc, a.f.rhi[j], s_b[((pair + 4) * SYN_NZ + slot) * SYN_PLANE + ixy]);
        }
      }
      s_z[e] = acc;
    }
    __syncthreads();
    for (int e = tid; e < 2 * SYN_TZ * TY * NX; e += 256) {
      const int ix = e % NX, r = e / NX;
      const int sy = r % TY, r2 = r / TY;
      const int sz = r2 % SYN_TZ, xb = r2 / SYN_TZ;
      float2 acc = make_float2(0.f, 0.f);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int idx = y0 + sy + 2 - j;
        if (idx >= 0 && idx < 2 * a.ncy && (idx & 1) == 0) {
          const int iy = (idx >> 1) - cy0;
          fma2(acc, a.f.rlo[j], s_z[((xb * SYN_TZ + sz) * NY + iy) * NX + ix]);
          fma2(acc, a.f.rhi[j], s_z[(((xb + 2) * SYN_TZ + sz) * NY + iy) * NX + ix]);
        }
      }
      s_y[e] = acc;
    }
    __syncthreads();
    for (int e = tid; e < SYN_TZ * TY * TX; e += 256) {
      const int sx = e % TX, r = e / TX;
      const int sy = r % TY, sz = r / TY;
      const int gx = x0 + sx, gy = y0 + sy, gz = z0 + sz;
      if (gx >= a.ax || gy >= a.ay || gz >= a.az) continue;
      float2 acc = make_float2(0.f, 0.f);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int idx = gx + 2 - j;
        if (idx >= 0 && idx < 2 * a.ncx && (idx & 1) == 0) {
          const int ix = (idx >> 1) - cx0;
          fma2(acc, a.f.rlo[j], s_y[((0 * SYN_TZ + sz) * TY + sy) * NX + ix]);
          fma2(acc, a.f.rhi[j], s_y[((1 * SYN_TZ + sz) * TY + sy) * NX + ix]);
        }
      }
      dst[(long long)gz * a.out.p2 + (long long)gy * a.out.p1 + gx] = acc;
    }
    // the ring slot the next step overwrites and s_z / s_y are free again
    __syncthreads();
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

// Level geometry: in_dims[l] (input box) and cdims[l] (coefficient box, 2nc per axis).
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

// Per-volume workspace: the band box of every level, 2*pad16(ncx) x 2ncy x 2ncz.
size_t wavelet_workspace_elems(const size_t* dims, int L, size_t* coef_elems, size_t* tmp_elems) {
  std::vector<std::array<size_t, 3>> in_dims, cdims;
  wavelet_geometry(dims, L, in_dims, cdims);
  size_t c = 0;
  for (int l = 0; l < L; ++l) c += 2 * pad16(cdims[l][0] / 2) * cdims[l][1] * cdims[l][2];
  *coef_elems = c;
  *tmp_elems = 0;
  return c;
}

// In-place wavelet soft threshold of nvols volumes; coef holds nvols *
// coef_elems elements (t1, t2 unused).
llr_status wavelet_denoise_device(llr_ctx* ctx, float2* vols, size_t nvols, const size_t* dims, float alpha, int L,
                                  float2* coef, float2* t1, float2* t2) {
  (void)t1;
  (void)t2;
  Db2F f = db2_filters();
  const size_t N = dims[0] * dims[1] * dims[2];
  if (L <= 0) return shrink_all_device(ctx, vols, N * nvols, alpha);  // the volume is its own approximation
  std::vector<std::array<size_t, 3>> in_dims, cdims;
  wavelet_geometry(dims, L, in_dims, cdims);
  std::vector<Box> band(L), input(L);
  std::vector<int> px(L);
  {
    float2* p = coef;
    for (int l = 0; l < L; ++l) {
      px[l] = (int)pad16(cdims[l][0] / 2);
      Box b;
      b.p = p;
      b.p1 = 2 * (long long)px[l];
      b.p2 = b.p1 * (long long)cdims[l][1];
      b.vp = b.p2 * (long long)cdims[l][2];
      band[l] = b;
      p += b.vp * nvols;
    }
    Box x0;
    x0.p = vols;
    x0.p1 = (long long)dims[0];
    x0.p2 = x0.p1 * (long long)dims[1];
    x0.vp = x0.p2 * (long long)dims[2];
    input[0] = x0;
    for (int l = 1; l < L; ++l) input[l] = band[l - 1];  // the approximation box sits at the origin
  }
  const size_t an_smem = (size_t)AN_SMEM * sizeof(float2);
  const size_t syn_smem = (size_t)SYN_SMEM * sizeof(float2);
  static bool attr_set = false;
  if (!attr_set) {
    LLR_HIP(hipFuncSetAttribute((const void*)k_wav_an3, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)an_smem));
    LLR_HIP(hipFuncSetAttribute((const void*)k_wav_syn3, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)syn_smem));
    attr_set = true;
  }
  hipStream_t s = ctx->stream;
  for (int l = 0; l < L; ++l) {
    const size_t* in = in_dims[l].data();
    const size_t* cd = cdims[l].data();
    AnArgs a{};
    a.in = input[l];
    a.ax = (int)in[0];
    a.ay = (int)in[1];
    a.az = (int)in[2];
    a.B = band[l];
    a.px = px[l];
    a.ncx = (int)cd[0] / 2;
    a.ncy = (int)cd[1] / 2;
    a.ncz = (int)cd[2] / 2;
    a.tiles_x = (a.ncx + TCX - 1) / TCX;
    a.tiles_y = (a.ncy + TCY - 1) / TCY;
    a.deepest = l == L - 1;
    a.alpha = alpha;
    a.f = f;
    dim3 g((unsigned)(a.tiles_x * a.tiles_y), (unsigned)nvols);
    hipLaunchKernelGGL(k_wav_an3, g, dim3(256), an_smem, s, a);
    LLR_HIP(hipGetLastError());
  }
  for (int l = L - 1; l >= 0; --l) {
    const size_t* od = in_dims[l].data();
    const size_t* cd = cdims[l].data();
    SynArgs a{};
    a.B = band[l];
    a.px = px[l];
    a.ncx = (int)cd[0] / 2;
    a.ncy = (int)cd[1] / 2;
    a.ncz = (int)cd[2] / 2;
    a.out = input[l];
    a.ax = (int)od[0];
    a.ay = (int)od[1];
    a.az = (int)od[2];
    a.tiles_x = (a.ax + TX - 1) / TX;
    a.tiles_y = (a.ay + TY - 1) / TY;
    a.f = f;
    dim3 g((unsigned)(a.tiles_x * a.tiles_y), (unsigned)nvols);
    hipLaunchKernelGGL(k_wav_syn3, g, dim3(256), syn_smem, s, a);
    LLR_HIP(hipGetLastError());
  }
  return LLR_OK;
}

}  // namespace llr
