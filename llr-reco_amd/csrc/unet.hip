// Warm-start 3D U-Net on gfx950: BF16 MFMA convolutions (FP32 accumulation)
// with hand-written forward and backward passes, Adam, and the packing of
// subspace coefficients into network channels.
//
// Tensors are (D, H, W, C) with channels fastest, C a multiple of 16.
// Activations are stored in BF16, gradients and the prediction in FP32.
// A k x k x k convolution (k = 3, zero padding 1; or k = 1) has FP32 master
// weights w[o][co][ci] (o = (dz*k + dy)*k + dx) and bias b[co]; two BF16
// copies are packed for the kernels: w[o][co][ci] (forward) and
// w[o][ci][co] (backward data, read at the mirrored offset).
//
// Convolutions are implicit GEMMs on v_mfma_f32_16x16x16_bf16: forward and
// backward-data take M = 16 voxels per wavefront, N = output channels
// (16 per MFMA tile), K = offsets x input channels (16 per step);
// backward-weight takes M = 16 output channels, N = 16 input channels and
// K = voxels, with per-wavefront partial sums added atomically.
//
// Network (levels L, widths c_l = base * 2^l): per level two 3^3 conv+ReLU,
// 2x2x2 average pooling between encoder levels; the decoder upsamples
// (nearest), concatenates [upsampled, skip] and applies two 3^3 conv+ReLU;
// a 1^3 output convolution is added to the input (residual prediction).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "common.h"
#include "internal.h"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16_t;

__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float((unsigned)h << 16); }
// round to nearest even (one v_cvt_pk_bf16_f32 on gfx950)
__device__ __forceinline__ bf16_t f2bf(float f) { return __builtin_bit_cast(bf16_t, (__bf16)f); }
__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// four consecutive channels as BF16 MFMA operand
__device__ __forceinline__ bf16x4 load4(const bf16_t* p) {
  uint2 u = *(const uint2*)p;
  bf16x4 r;
  r[0] = (short)(u.x & 0xFFFF);
  r[1] = (short)(u.x >> 16);
  r[2] = (short)(u.y & 0xFFFF);
  r[3] = (short)(u.y >> 16);
  return r;
}
__device__ __forceinline__ bf16x4 load4(const float* p) {
  float4 f = *(const float4*)p;
  bf16x4 r;
  r[0] = (short)f2bf(f.x);
  r[1] = (short)f2bf(f.y);
  r[2] = (short)f2bf(f.z);
  r[3] = (short)f2bf(f.w);
  return r;
}

constexpr int MAX_CO_TILES = 16;  // up to 256 output channels
constexpr int MAX_MIX = 8;        // subspace coefficients mapped to network channels

struct ConvArgs {
  const void* in;       // D*H*W x cin (BF16 activations or FP32 gradients)
  const bf16_t* w;      // [kk][cout][cin]
  const float* bias;    // cout or null
  const bf16_t* mask;   // D*H*W x cout: output *= (mask > 0), or null
  void* out;            // D*H*W x cout (BF16 or FP32)
  int D, H, W, cin, cout, k;
  int mirror;           // read offset kk-1-o (backward data)
  int relu, accumulate;
  const bf16_t* up;     // 3^3 BF16 forward only: half-resolution source of input channels [0, cup), or null
  int cup;
  const float* addend;  // 1^3 forward only: out = conv + addend (same layout as out), or null
  int per_xcd;          // 3^3: voxel runs per XCD (grid.x = 8 per_xcd; see k_conv3)
};

// out[v][co] = sum_{o, ci} in[v + off(o)][ci] w[o][co][ci] (+ bias, ReLU, mask, +=)
template <typename TIn, typename TOut>
__global__ void __launch_bounds__(256) k_conv(ConvArgs a) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int n = lane & 15, kq = lane >> 4;
  const long long V = (long long)a.D * a.H * a.W;
  const long long v0 = ((long long)blockIdx.x * 4 + wave) * 16;
  if (v0 >= V) return;
  const int nt = a.cout / 16;
  const int kk = a.k * a.k * a.k, p = a.k / 2;
  // the A-operand voxel of this lane
  const long long va = v0 + n;
  int x = 0, y = 0, z = 0;
  if (va < V) {
    x = (int)(va % a.W);
    y = (int)((va / a.W) % a.H);
    z = (int)(va / ((long long)a.W * a.H));
  }
  f32x4 acc[MAX_CO_TILES];
#pragma unroll
  for (int t = 0; t < MAX_CO_TILES; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const TIn* in = (const TIn*)a.in;
  for (int o = 0; o < kk; ++o) {
    const int dz = o / (a.k * a.k) - p, dy = (o / a.k) % a.k - p, dx = o % a.k - p;
    const int zz = z + dz, yy = y + dy, xx = x + dx;
    const bool valid = va < V && zz >= 0 && zz < a.D && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W;
    const long long src = (((long long)zz * a.H + yy) * a.W + xx) * a.cin;
    const int ow = a.mirror ? kk - 1 - o : o;
    const bf16_t* wo = a.w + (size_t)ow * a.cout * a.cin;
    for (int c0 = 0; c0 < a.cin; c0 += 16) {
      bf16x4 av = {0, 0, 0, 0};
      if (valid) av = load4(in + src + c0 + 4 * kq);
#pragma unroll
      for (int t = 0; t < MAX_CO_TILES; ++t) {
        if (t < nt) {
          bf16x4 bv = load4(wo + (size_t)(16 * t + n) * a.cin + c0 + 4 * kq);
          acc[t] = mfma16(av, bv, acc[t]);
        }
      }
    }
  }
  TOut* out = (TOut*)a.out;
#pragma unroll
  for (int t = 0; t < MAX_CO_TILES; ++t) {
    if (t < nt) {
      const int co = 16 * t + n;
      const float b = a.bias ? a.bias[co] : 0.f;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const long long v = v0 + 4 * kq + j;
        if (v < V) {
          float r = acc[t][j] + b;
          if (a.relu) r = fmaxf(r, 0.f);
          const size_t idx = (size_t)v * a.cout + co;
          if (a.mask && !(bf2f(a.mask[idx]) > 0.f)) r = 0.f;
          if constexpr (sizeof(TOut) == 4) {
            if (a.accumulate) r += ((float*)out)[idx];
            if (a.addend) r += a.addend[idx];
            ((float*)out)[idx] = r;
          } else {
            ((bf16_t*)out)[idx] = f2bf(r);
          }
        }
      }
    }
  }
}

// 3^3 convolution with the input and the weights staged in LDS. A workgroup
// owns CV_M consecutive voxels (each wavefront CV_MT m-tiles of 16); for every
// chunk of 16 input channels it stages the nine (dz, dy) runs of CV_M + 2
// voxels its 27 offsets read, so each input value is fetched from memory about
// 9 / 3 times per workgroup instead of once per offset and lane, and the
// chunk's 27 x cout x 16 weights, so the MFMA loop reads only LDS. Rows are
// CV_ROW = 16 values (32 bytes, four 8-byte pieces). An operand read takes
// piece kq of 16 consecutive rows; the LDS serves it 16 lanes at a time, and
// over its 32 banks rows n, n+4, n+8, n+12 would share banks. So row R keeps
// piece p at slot p ^ key(R), key(R) = (R' >> 2) & 3 with R' the row's
// position in its run (or block of 16 weight rows): those four rows then use
// four different slots and the read is conflict-free. (Unswizzled, the
// counters read 18 stall cycles per LDS instruction; padding rows to 40 bytes
// removes them too but costs 13 KB of LDS and a resident workgroup.)
// FP32 inputs (the backward-data gradients) are rounded to
// BF16 while staged. With a.up set, input channels [0, a.cup) are the nearest
// upsampling of the half-resolution tensor a.up and channels [a.cup, cin) come
// from a.in (the decoder's [upsampled, skip] concatenation, never stored).
// Neighbours outside the volume: with ROWS (the row length W a multiple of
// CV_M, so a run lies inside one x-row) every staged row is inside or outside
// the volume as a whole and only the run's two end values can fall off the
// row, so the staging writes zeros there and the MFMA loop reads LDS without
// masks; otherwise they are masked per lane, as in k_conv.
constexpr int CV_MT = 2;
constexpr int CV_M = 64 * CV_MT;
constexpr int CV_RUN = CV_M + 2;
constexpr int CV_ROW = 16;

__device__ __forceinline__ bf16x4 lds4(const bf16_t* p) { return __builtin_bit_cast(bf16x4, *(const uint2*)p); }

template <typename TIn, typename TOut, int NTM, bool ROWS>
__global__ void __launch_bounds__(256) k_conv3(ConvArgs a) {
  __shared__ __align__(16) bf16_t s_run[9 * CV_RUN * CV_ROW];
  __shared__ __align__(16) bf16_t s_w[27 * 16 * NTM * CV_ROW];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int n = lane & 15, kq = lane >> 4;
  // voxel indices are 32-bit (the host checks D*H*W*channels < 2^31)
  const int HW = a.H * a.W, V = HW * a.D;
  // workgroups are dealt round-robin to the 8 XCDs; XCD j takes the contiguous
  // runs [j per_xcd, (j + 1) per_xcd), so the (dz, dy) neighbour rows a run
  // stages were mostly fetched into the same L2 by the runs before it
  const int run = (int)(blockIdx.x % 8) * a.per_xcd + (int)(blockIdx.x / 8);
  if (run * CV_M >= V) return;
  const int vb = run * CV_M;
  const int cb = blockIdx.y * 16 * NTM;             // this workgroup's output channels [cb, cb + 16 nt)
  const int nt = min(NTM, (a.cout - cb) / 16);
  const int cin2 = a.cin - a.cup;  // channels per voxel of a.in
  // per m-tile lane voxel: the 27-bit set of offsets o whose neighbour lies inside the volume
  unsigned vmask[CV_MT];
#pragma unroll
  for (int mt = 0; mt < CV_MT; ++mt) {
    vmask[mt] = 0u;
    if constexpr (ROWS) continue;
    const int v = vb + (wave * CV_MT + mt) * 16 + n;
    const int x = v % a.W, y = (v / a.W) % a.H, z = v / HW;
    // bit d + 1 of an axis mask: the step d in {-1, 0, 1} stays inside
    const unsigned mx = (x > 0 ? 1u : 0u) | 2u | (x < a.W - 1 ? 4u : 0u);
    const unsigned my = (y > 0 ? 1u : 0u) | 2u | (y < a.H - 1 ? 4u : 0u);
    const unsigned mz = (z > 0 ? 1u : 0u) | 2u | (z < a.D - 1 ? 4u : 0u);
    unsigned m = 0;
#pragma unroll
    for (int o = 0; o < 27; ++o) m |= ((mz >> (o / 9)) & (my >> ((o / 3) % 3)) & (mx >> (o % 3)) & 1u) << o;
    vmask[mt] = v < V ? m : 0u;
  }
  f32x4 acc[CV_MT][NTM];
#pragma unroll
  for (int mt = 0; mt < CV_MT; ++mt)
#pragma unroll
    for (int t = 0; t < NTM; ++t) acc[mt][t] = f32x4{0.f, 0.f, 0.f, 0.f};
  // every load of a chunk is issued before the first LDS store (the loads
  // are independent; issued one at a time their latencies would add up)
  constexpr int NRL = (9 * CV_RUN * 2 + 255) / 256;
  constexpr int NWL = (27 * 16 * NTM * 2 + 255) / 256;
  // ROWS: the run's row (x0 .. x0 + CV_M - 1, y0, z0), found once
  const int yz0 = ROWS ? vb / a.W : 0;
  const int x0 = vb - yz0 * a.W, y0 = ROWS ? yz0 % a.H : 0, z0 = ROWS ? yz0 / a.H : 0;
  // ROWS staging: thread t takes run position t >> 1 (channel half t & 1) of
  // each of the nine rows, so a row's (y, z) part of the address and its
  // inside test are uniform; positions CV_M, CV_M + 1 of the nine rows (36
  // values) go to threads 0..35 as load NRL - 1
  static_assert(!ROWS || (CV_M == 128 && NRL == 10), "ROWS staging: 256 threads over 128 positions, 9 + 1 loads");
  auto stage_rows = [&](int q, int& r, int& i, int& half) {
    if (q < 9) {
      r = q;
      i = tid >> 1;
      half = tid & 1;
    } else {
      r = tid >> 2;
      i = CV_M + ((tid >> 1) & 1);
      half = tid & 1;
    }
  };
  for (int c0 = 0; c0 < a.cin; c0 += 16) {
    const bool from_up = a.up != nullptr && c0 < a.cup;
    uint4 ur[NRL];
    // branch-free: every load from an address clamped into the tensor and
    // zeroed afterwards when its neighbour is outside (a load under a branch
    // was waited for at once, so the chunk's loads ran one at a time)
    bool rin[NRL];
    if constexpr (ROWS) {
#pragma unroll
      for (int q = 0; q < NRL; ++q) {
        int r, i, half;
        stage_rows(q, r, i, half);
        r = min(r, 8);
        const int xx = x0 - 1 + i, yy = y0 + r % 3 - 1, zz = z0 + r / 3 - 1;
        rin[q] = (q < NRL - 1 || tid < 36) && xx >= 0 && xx < a.W && yy >= 0 && yy < a.H && zz >= 0 && zz < a.D;
        const int xc = min(max(xx, 0), a.W - 1), yc = min(max(yy, 0), a.H - 1), zc = min(max(zz, 0), a.D - 1);
        if constexpr (sizeof(TIn) == 2) {
          const int ci = ((zc >> 1) * (a.H >> 1) + (yc >> 1)) * (a.W >> 1) + (xc >> 1);
          const int idx = (zc * a.H + yc) * a.W + xc;
          const uint4* up = (const uint4*)(a.up + ci * a.cup + c0 + 8 * half);
          const uint4* ip = (const uint4*)((const bf16_t*)a.in + idx * cin2 + (c0 - a.cup) + 8 * half);
          ur[q] = *(from_up ? up : ip);
        } else {
          const int idx = (zc * a.H + yc) * a.W + xc;
          const float* fp = (const float*)a.in + idx * a.cin + c0 + 8 * half;
          const float4 f0 = *(const float4*)fp, f1 = *(const float4*)(fp + 4);
          uint4& u = ur[q];
          u.x = (unsigned)f2bf(f0.x) | ((unsigned)f2bf(f0.y) << 16);
          u.y = (unsigned)f2bf(f0.z) | ((unsigned)f2bf(f0.w) << 16);
          u.z = (unsigned)f2bf(f1.x) | ((unsigned)f2bf(f1.y) << 16);
          u.w = (unsigned)f2bf(f1.z) | ((unsigned)f2bf(f1.w) << 16);
        }
      }
    } else {
      // the same, branch-free: the flat neighbour index clamped into the
      // volume, zeroed afterwards when outside (the x / y wraps of a flat
      // index are masked per lane in the MFMA loop)
#pragma unroll
      for (int q = 0; q < NRL; ++q) {
        const int e = tid + 256 * q;
        const int half = e & 1, i = min((e >> 1) % CV_RUN, CV_RUN - 1), r = min((e >> 1) / CV_RUN, 8);
        const int idx = vb - 1 + i + (r / 3 - 1) * HW + (r % 3 - 1) * a.W;
        rin[q] = e < 9 * CV_RUN * 2 && idx >= 0 && idx < V;
        const int ic = min(max(idx, 0), V - 1);
        if constexpr (sizeof(TIn) == 2) {
          const int x = ic % a.W, y = (ic / a.W) % a.H, z = ic / HW;
          const int ci = ((z >> 1) * (a.H >> 1) + (y >> 1)) * (a.W >> 1) + (x >> 1);
          const uint4* up = (const uint4*)(a.up + ci * a.cup + c0 + 8 * half);
          const uint4* ip = (const uint4*)((const bf16_t*)a.in + ic * cin2 + (c0 - a.cup) + 8 * half);
          ur[q] = *(from_up ? up : ip);
        } else {
          const float* fp = (const float*)a.in + ic * a.cin + c0 + 8 * half;
          const float4 f0 = *(const float4*)fp, f1 = *(const float4*)(fp + 4);
          uint4& u = ur[q];
          u.x = (unsigned)f2bf(f0.x) | ((unsigned)f2bf(f0.y) << 16);
          u.y = (unsigned)f2bf(f0.z) | ((unsigned)f2bf(f0.w) << 16);
          u.z = (unsigned)f2bf(f1.x) | ((unsigned)f2bf(f1.y) << 16);
          u.w = (unsigned)f2bf(f1.z) | ((unsigned)f2bf(f1.w) << 16);
        }
      }
    }
    // this chunk's weights: s_w[o][co - cb][0..16) = w[o'][co][c0..c0+16), o' = o or 26 - o
    // (zeros for the tiles t >= nt, so every tile's MFMAs run without a branch)
    const int ncw = 16 * nt;
    uint4 uw[NWL];
#pragma unroll
    for (int q = 0; q < NWL; ++q) {
      const int e = tid + 256 * q;
      uw[q] = make_uint4(0u, 0u, 0u, 0u);
      if (e >= 27 * 16 * NTM * 2) continue;
      const int half = e & 1, col = (e >> 1) % (16 * NTM), o = (e >> 1) / (16 * NTM);
      const int ow = a.mirror ? 26 - o : o;
      if (col < ncw) uw[q] = *(const uint4*)(a.w + ((size_t)ow * a.cout + cb + col) * a.cin + c0 + 8 * half);
    }
#pragma unroll
    for (int q = 0; q < NRL; ++q) {
      int half, i, r;
      if constexpr (ROWS) {
        if (q == NRL - 1 && tid >= 36) continue;
        stage_rows(q, r, i, half);
      } else {
        const int e = tid + 256 * q;
        if (e >= 9 * CV_RUN * 2) continue;
        half = e & 1;
        i = (e >> 1) % CV_RUN;
        r = (e >> 1) / CV_RUN;
      }
      const int key = (i >> 2) & 3;  // pieces 2 half, 2 half + 1 -> slots (2 half + {0, 1}) ^ key
      uint4 v = (key & 1) ? make_uint4(ur[q].z, ur[q].w, ur[q].x, ur[q].y) : ur[q];
      if (!rin[q]) v = make_uint4(0u, 0u, 0u, 0u);
      *(uint4*)(s_run + (size_t)(r * CV_RUN + i) * CV_ROW + 8 * (half ^ (key >> 1))) = v;
    }
#pragma unroll
    for (int q = 0; q < NWL; ++q) {
      const int e = tid + 256 * q;
      if (e >= 27 * 16 * NTM * 2) continue;
      const int half = e & 1, col = (e >> 1) % (16 * NTM), o = (e >> 1) / (16 * NTM);
      const int key = (col >> 2) & 3;
      const uint4 v = (key & 1) ? make_uint4(uw[q].z, uw[q].w, uw[q].x, uw[q].y) : uw[q];
      *(uint4*)(s_w + (size_t)(o * 16 * NTM + col) * CV_ROW + 8 * (half ^ (key >> 1))) = v;
    }
    __syncthreads();
#pragma unroll
    for (int o = 0; o < 27; ++o) {
      const int dx = o % 3 - 1, r = o / 3;
      // the swizzled slot of piece kq: pos = 16 (wave CV_MT + mt) + n + dx + 1, so
      // (pos >> 2) & 3 = ((n + dx + 1) >> 2) & 3 for every m-tile
      const int pa = kq ^ (((n + dx + 1) >> 2) & 3);
      bf16x4 av[CV_MT];
#pragma unroll
      for (int mt = 0; mt < CV_MT; ++mt) {
        const int pos = (wave * CV_MT + mt) * 16 + n + dx + 1;
        if constexpr (ROWS) {
          av[mt] = lds4(s_run + (size_t)(r * CV_RUN + pos) * CV_ROW + 4 * pa);
        } else {
          const bool valid = (vmask[mt] >> o) & 1u;
          av[mt] = bf16x4{0, 0, 0, 0};
          if (valid) av[mt] = lds4(s_run + (size_t)(r * CV_RUN + pos) * CV_ROW + 4 * pa);
        }
      }
      const bf16_t* wo = s_w + (size_t)o * 16 * NTM * CV_ROW;
#pragma unroll
      for (int t = 0; t < NTM; ++t) {
        const bf16x4 bv = lds4(wo + (size_t)(16 * t + n) * CV_ROW + 4 * (kq ^ ((n >> 2) & 3)));
#pragma unroll
        for (int mt = 0; mt < CV_MT; ++mt) acc[mt][t] = mfma16(av[mt], bv, acc[mt][t]);
      }
    }
    __syncthreads();
  }
  TOut* out = (TOut*)a.out;
#pragma unroll
  for (int mt = 0; mt < CV_MT; ++mt) {
#pragma unroll
    for (int t = 0; t < NTM; ++t) {
      if (t < nt) {
        const int co = cb + 16 * t + n;
        const float b = a.bias ? a.bias[co] : 0.f;
        // 32-bit indices (the host checks D H W cout < 2^31); a row-aligned run
        // lies inside the volume
        const int v0 = vb + (wave * CV_MT + mt) * 16 + 4 * kq;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int v = v0 + j;
          if (ROWS || v < V) {
            float rr = acc[mt][t][j] + b;
            if (a.relu) rr = fmaxf(rr, 0.f);
            const int idx = v * a.cout + co;
            if (a.mask && !(bf2f(a.mask[idx]) > 0.f)) rr = 0.f;
            if constexpr (sizeof(TOut) == 4) {
              if (a.accumulate) rr += ((float*)out)[idx];
              ((float*)out)[idx] = rr;
            } else {
              ((bf16_t*)out)[idx] = f2bf(rr);
            }
          }
        }
      }
    }
  }
}

// 1^3 convolution with 16 input and 16 output channels into FP32 (the output
// layer): out[v][co] = sum_ci w[co][ci] in[v][ci] + bias[co] (ReLU if set),
// plus addend[v][co] when given. One thread per voxel: two 16-byte loads of
// its BF16 channels, the weights (FP32 copies in LDS, read by every lane at
// the same address) and FP32 FMAs (products of BF16 values are exact in FP32,
// as in the MFMA form; only the summation order differs), four 16-byte loads
// of the addend and four 16-byte stores. k_conv's 16-voxel wavefront tiles
// store 4 bytes per lane: 0.87 ms at 256x256x128.
__global__ void __launch_bounds__(256) k_conv1_16(const bf16_t* __restrict__ in, const bf16_t* __restrict__ w,
                                                  const float* __restrict__ bias, const float* __restrict__ addend,
                                                  float* __restrict__ out, long long V, int relu) {
  __shared__ float4 sw[16 * 4];  // [co][ci / 4]
  __shared__ float sb[16];
  const int tid = threadIdx.x;
  if (tid < 64) {
    const uint2 u = *(const uint2*)(w + 4 * tid);
    sw[tid] = make_float4(bf2f((bf16_t)(u.x & 0xFFFF)), bf2f((bf16_t)(u.x >> 16)), bf2f((bf16_t)(u.y & 0xFFFF)),
                          bf2f((bf16_t)(u.y >> 16)));
  }
  if (tid < 16) sb[tid] = bias ? bias[tid] : 0.f;
  __syncthreads();
  for (long long v = blockIdx.x * (long long)blockDim.x + tid; v < V; v += (long long)gridDim.x * blockDim.x) {
    const uint4* p = (const uint4*)(in + (size_t)v * 16);
    const uint4 u0 = p[0], u1 = p[1];
    const unsigned wd[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
    float x[16];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      x[2 * i] = __uint_as_float(wd[i] << 16);
      x[2 * i + 1] = __uint_as_float(wd[i] & 0xFFFF0000u);
    }
    float r[16];
#pragma unroll
    for (int co = 0; co < 16; ++co) {
      // the row index passes through an empty asm statement, so the weights are
      // read here, row by row, and not hoisted out of the voxel loop into 256
      // registers
      int wi = co * 4;
      asm volatile("" : "+v"(wi));
      float acc = sb[co];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float4 wv = sw[wi + q];
        acc = fmaf(wv.x, x[4 * q], acc);
        acc = fmaf(wv.y, x[4 * q + 1], acc);
        acc = fmaf(wv.z, x[4 * q + 2], acc);
        acc = fmaf(wv.w, x[4 * q + 3], acc);
      }
      r[co] = relu ? fmaxf(acc, 0.f) : acc;
    }
    float4* o = (float4*)(out + (size_t)v * 16);
    if (addend) {
      const float4* ad = (const float4*)(addend + (size_t)v * 16);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float4 a4 = ad[q];
        r[4 * q] += a4.x;
        r[4 * q + 1] += a4.y;
        r[4 * q + 2] += a4.z;
        r[4 * q + 3] += a4.w;
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) o[q] = make_float4(r[4 * q], r[4 * q + 1], r[4 * q + 2], r[4 * q + 3]);
  }
}

// dW[o][co][ci] += sum_v g[v][co] x[v + off(o)][ci] over one chunk of voxels
// per wavefront (M = 16 co, N = 16 ci, K = voxels), for all KK offsets at
// once: the g operand of 16 voxels is loaded once and feeds KK MFMAs.
template <int KK>
__global__ void __launch_bounds__(64) k_conv_wgrad(const float* __restrict__ g, const bf16_t* __restrict__ x,
                                                   float* __restrict__ dw, int D, int H, int W, int cin, int cout,
                                                   long long chunk) {
  constexpr int k = KK == 27 ? 3 : 1, p = k / 2;
  const int lane = threadIdx.x;
  const int n = lane & 15, kq = lane >> 4;
  const int tci = cin / 16;
  const int ti = blockIdx.y % tci, tc = blockIdx.y / tci;
  const long long V = (long long)D * H * W;
  const long long lo = (long long)blockIdx.x * chunk, hi = min(V, lo + chunk);
  f32x4 acc[KK];
#pragma unroll
  for (int o = 0; o < KK; ++o) acc[o] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (long long v0 = lo; v0 < hi; v0 += 16) {
    bf16x4 av;
    int vx[4], vy[4], vz[4];
    bool in[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const long long v = v0 + 4 * kq + j;
      in[j] = v < hi;
      float gv = 0.f;
      vx[j] = vy[j] = vz[j] = 0;
      if (in[j]) {
        gv = g[(size_t)v * cout + 16 * tc + n];
        vx[j] = (int)(v % W);
        vy[j] = (int)((v / W) % H);
        vz[j] = (int)(v / ((long long)W * H));
      }
      av[j] = (short)f2bf(gv);
    }
#pragma unroll
    for (int o = 0; o < KK; ++o) {
      const int dz = o / (k * k) - p, dy = (o / k) % k - p, dx = o % k - p;
      bf16x4 bv;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int xx = vx[j] + dx, yy = vy[j] + dy, zz = vz[j] + dz;
        bf16_t xv = 0;
        if (in[j] && xx >= 0 && xx < W && yy >= 0 && yy < H && zz >= 0 && zz < D)
          xv = x[(((size_t)zz * H + yy) * W + xx) * cin + 16 * ti + n];
        bv[j] = (short)xv;
      }
      acc[o] = mfma16(av, bv, acc[o]);
    }
  }
#pragma unroll
  for (int o = 0; o < KK; ++o) {
    float* dst = dw + (size_t)o * cout * cin;
#pragma unroll
    for (int j = 0; j < 4; ++j) atomicAdd(dst + (size_t)(16 * tc + 4 * kq + j) * cin + 16 * ti + n, acc[o][j]);
  }
}

// db[c] += sum_v g[v][c]: per-thread sums over rows of the block, an LDS
// reduction over the block's rows, then one atomic per channel per block
__global__ void __launch_bounds__(256) k_bias_grad(const float* __restrict__ g, float* __restrict__ db, long long V,
                                                   int c) {
  __shared__ float sh[256];
  const int ch = threadIdx.x % c;  // c <= 256, a multiple of 16
  const int rows = 256 / c;
  const int row = threadIdx.x / c;
  float s = 0.f;
  if (row < rows)
    for (long long v = (long long)blockIdx.x * rows + row; v < V; v += (long long)gridDim.x * rows)
      s += g[(size_t)v * c + ch];
  sh[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < c) {
    float t = 0.f;
    for (int r = 0; r < rows; ++r) t += sh[r * c + threadIdx.x];
    atomicAdd(db + threadIdx.x, t);
  }
}

// 2x2x2 average pooling (BF16 -> BF16)
__global__ void k_pool_fwd(const bf16_t* __restrict__ in, bf16_t* __restrict__ out, int D, int H, int W, int C) {
  const int d = D / 2, h = H / 2, w = W / 2;
  const long long total = (long long)d * h * w * C;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x) {
    int c = (int)(e % C);
    long long v = e / C;
    int x = (int)(v % w), y = (int)((v / w) % h), z = (int)(v / ((long long)w * h));
    float s = 0.f;
    for (int q = 0; q < 8; ++q)
      s += bf2f(in[((((size_t)(2 * z + (q >> 2)) * H + 2 * y + ((q >> 1) & 1)) * W) + 2 * x + (q & 1)) * C + c]);
    out[e] = f2bf(s * 0.125f);
  }
}

// dst_fine = (dst_fine + g_coarse[parent] / 8) * (mask > 0)
__global__ void k_pool_bwd_add(const float* __restrict__ g, float* __restrict__ dst, const bf16_t* __restrict__ mask,
                               int D, int H, int W, int C) {
  const long long total = (long long)D * H * W * C;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x) {
    int c = (int)(e % C);
    long long v = e / C;
    int x = (int)(v % W), y = (int)((v / W) % H), z = (int)(v / ((long long)W * H));
    float r = dst[e] + 0.125f * g[((((size_t)(z / 2) * (H / 2)) + y / 2) * (W / 2) + x / 2) * C + c];
    if (mask && !(bf2f(mask[e]) > 0.f)) r = 0.f;
    dst[e] = r;
  }
}

// concat buffer (Ct channels): [0, Cu) = nearest upsampling of the coarse
// tensor, [Cu, Cu + Cs) = the skip tensor
__global__ void k_concat_fwd(const bf16_t* __restrict__ coarse, const bf16_t* __restrict__ skip,
                             bf16_t* __restrict__ out, int D, int H, int W, int Cu, int Cs) {
  const int Ct = Cu + Cs;
  const long long total = (long long)D * H * W * Ct;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x) {
    int c = (int)(e % Ct);
    long long v = e / Ct;
    if (c < Cu) {
      int x = (int)(v % W), y = (int)((v / W) % H), z = (int)(v / ((long long)W * H));
      out[e] = coarse[((((size_t)(z / 2) * (H / 2)) + y / 2) * (W / 2) + x / 2) * Cu + c];
    } else {
      out[e] = skip[(size_t)v * Cs + (c - Cu)];
    }
  }
}

// backward of the concat: coarse gradient = sum over the 8 children of the
// upsampled channels (times the coarse activation's ReLU mask); skip gradient
// = the remaining channels
__global__ void k_concat_bwd(const float* __restrict__ g, float* __restrict__ g_coarse,
                             const bf16_t* __restrict__ coarse_mask, float* __restrict__ g_skip, int D, int H, int W,
                             int Cu, int Cs) {
  const int Ct = Cu + Cs;
  const int d = D / 2, h = H / 2, w = W / 2;
  const long long nc = (long long)d * h * w * Cu;
  const long long ns = (long long)D * H * W * Cs;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < nc + ns; e += (long long)gridDim.x * blockDim.x) {
    if (e < nc) {
      int c = (int)(e % Cu);
      long long v = e / Cu;
      int x = (int)(v % w), y = (int)((v / w) % h), z = (int)(v / ((long long)w * h));
      float s = 0.f;
      for (int q = 0; q < 8; ++q)
        s += g[((((size_t)(2 * z + (q >> 2)) * H + 2 * y + ((q >> 1) & 1)) * W) + 2 * x + (q & 1)) * Ct + c];
      if (coarse_mask && !(bf2f(coarse_mask[e]) > 0.f)) s = 0.f;
      g_coarse[e] = s;
    } else {
      long long f = e - nc;
      int c = (int)(f % Cs);
      long long v = f / Cs;
      g_skip[f] = g[(size_t)v * Ct + Cu + c];
    }
  }
}

// grad = 2 (pred - target) / n; partial sums of (pred - target)^2 per block
__global__ void __launch_bounds__(256) k_mse(const float* __restrict__ pred, const float* __restrict__ tgt,
                                             float* __restrict__ grad, long long n, double* partial) {
  __shared__ double sh[4];
  double s = 0.0;
  const float sc = 2.0f / (float)n;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
    float d = pred[e] - tgt[e];
    grad[e] = sc * d;
    s += (double)d * d;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// Adam (Kingma & Ba, Alg. 1) with bias correction folded into the step size
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       long long n, float lr_t, float b1, float b2, float eps_t) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
    float gi = g[e];
    float mi = b1 * m[e] + (1.f - b1) * gi;
    float vi = b2 * v[e] + (1.f - b2) * gi * gi;
    m[e] = mi;
    v[e] = vi;
    p[e] -= lr_t * mi / (sqrtf(vi) + eps_t);
  }
}

// BF16 forward [o][co][ci] and transposed [o][ci][co] copies of one layer
__global__ void k_pack_weights(const float* __restrict__ w, bf16_t* __restrict__ wf, bf16_t* __restrict__ wt, int kk,
                               int cout, int cin) {
  const long long total = (long long)kk * cout * cin;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x) {
    int ci = (int)(e % cin);
    long long r = e / cin;
    int co = (int)(r % cout);
    int o = (int)(r / cout);
    bf16_t h = f2bf(w[e]);
    wf[e] = h;
    wt[((size_t)o * cin + ci) * cout + co] = h;
  }
}

__global__ void k_f32_to_bf16(const float* __restrict__ a, bf16_t* __restrict__ b, long long n) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
    b[e] = f2bf(a[e]);
}

// Real K x K mixing of complex coefficients (row-major), passed by value.
struct Mix {
  float m[MAX_MIX * MAX_MIX];
};

// K complex planes of N voxels (x fastest) -> (N, C) channels
// [Re b0, Im b0, Re b1, ...] / scale with b = M c per voxel, remaining
// channels zero. One thread per voxel: K coalesced plane reads, C/4 16-byte
// stores (FP32) or C/8 (BF16). outb (FP32 out only, or null): the same
// channels rounded to BF16 as well (the network's input beside the residual).
template <typename TOut>
__global__ void __launch_bounds__(256) k_pack_coef(const float2* __restrict__ c, TOut* __restrict__ out, long long N,
                                                   int K, int C, Mix M, float inv_scale,
                                                   bf16_t* __restrict__ outb = nullptr) {
  for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < N; v += (long long)gridDim.x * blockDim.x) {
    float2 z[MAX_MIX];
#pragma unroll
    for (int k = 0; k < MAX_MIX; ++k) z[k] = k < K ? c[(size_t)k * N + v] : make_float2(0.f, 0.f);
    float ch[2 * MAX_MIX];
#pragma unroll
    for (int j = 0; j < MAX_MIX; ++j) {
      float2 b = make_float2(0.f, 0.f);
#pragma unroll
      for (int k = 0; k < MAX_MIX; ++k) {
        b.x += M.m[j * MAX_MIX + k] * z[k].x;
        b.y += M.m[j * MAX_MIX + k] * z[k].y;
      }
      ch[2 * j] = j < K ? b.x * inv_scale : 0.f;
      ch[2 * j + 1] = j < K ? b.y * inv_scale : 0.f;
    }
    TOut* o = out + (size_t)v * C;  // C >= 2 * MAX_MIX, a multiple of 16
    if constexpr (sizeof(TOut) == 4) {
      if (out) {  // null: the BF16 channels only
#pragma unroll
        for (int q = 0; q < 2 * MAX_MIX; q += 4)
          *(float4*)(o + q) = make_float4(ch[q], ch[q + 1], ch[q + 2], ch[q + 3]);
        for (int q = 2 * MAX_MIX; q < C; q += 4) *(float4*)(o + q) = make_float4(0.f, 0.f, 0.f, 0.f);
      }
      if (outb) {
        bf16_t* ob = outb + (size_t)v * C;
#pragma unroll
        for (int q = 0; q < 2 * MAX_MIX; q += 8) {
          uint4 u;
          u.x = (unsigned)f2bf(ch[q]) | ((unsigned)f2bf(ch[q + 1]) << 16);
          u.y = (unsigned)f2bf(ch[q + 2]) | ((unsigned)f2bf(ch[q + 3]) << 16);
          u.z = (unsigned)f2bf(ch[q + 4]) | ((unsigned)f2bf(ch[q + 5]) << 16);
          u.w = (unsigned)f2bf(ch[q + 6]) | ((unsigned)f2bf(ch[q + 7]) << 16);
          *(uint4*)(ob + q) = u;
        }
        for (int q = 2 * MAX_MIX; q < C; q += 8) *(uint4*)(ob + q) = make_uint4(0u, 0u, 0u, 0u);
      }
    } else {
#pragma unroll
      for (int q = 0; q < 2 * MAX_MIX; q += 8) {
        uint4 u;
        u.x = (unsigned)f2bf(ch[q]) | ((unsigned)f2bf(ch[q + 1]) << 16);
        u.y = (unsigned)f2bf(ch[q + 2]) | ((unsigned)f2bf(ch[q + 3]) << 16);
        u.z = (unsigned)f2bf(ch[q + 4]) | ((unsigned)f2bf(ch[q + 5]) << 16);
        u.w = (unsigned)f2bf(ch[q + 6]) | ((unsigned)f2bf(ch[q + 7]) << 16);
        *(uint4*)(o + q) = u;
      }
      for (int q = 2 * MAX_MIX; q < C; q += 8) *(uint4*)(o + q) = make_uint4(0u, 0u, 0u, 0u);
    }
  }
}

// (N, C) FP32 channels -> K complex planes: c = M b * scale per voxel
__global__ void __launch_bounds__(256) k_unpack_coef(const float* __restrict__ in, float2* __restrict__ c, long long N,
                                                     int K, int C, Mix M, float scale) {
  for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < N; v += (long long)gridDim.x * blockDim.x) {
    const float* p = in + (size_t)v * C;
    float2 b[MAX_MIX];
#pragma unroll
    for (int j = 0; j < MAX_MIX; ++j) b[j] = j < K ? make_float2(p[2 * j], p[2 * j + 1]) : make_float2(0.f, 0.f);
#pragma unroll
    for (int k = 0; k < MAX_MIX; ++k) {
      if (k < K) {
        float2 z = make_float2(0.f, 0.f);
#pragma unroll
        for (int j = 0; j < MAX_MIX; ++j) {
          z.x += M.m[k * MAX_MIX + j] * b[j].x;
          z.y += M.m[k * MAX_MIX + j] * b[j].y;
        }
        c[(size_t)k * N + v] = make_float2(z.x * scale, z.y * scale);
      }
    }
  }
}

// The output layer at inference (1^3 convolution, 16 -> 16 channels, as
// k_conv1_16) fused with the unpacking: per voxel r = W x + bias, then
// out_k = c_k + scale sum_j M[k][j] (r_2j, r_2j+1) for k < K. The residual is
// the input coefficient itself: the packing's channel mixing and M are exact
// inverses on the subspace, so the unpacked packed input is c up to rounding.
// Neither the FP32 input channels nor the FP32 prediction is stored. out may
// be c (each voxel is read before it is written, by the same thread).
__global__ void __launch_bounds__(256) k_out_coef(const bf16_t* __restrict__ in, const bf16_t* __restrict__ w,
                                                  const float* __restrict__ bias, const float2* c, float2* out,
                                                  long long N, int K, Mix M, float scale, int relu) {
  __shared__ float4 sw[16 * 4];  // [co][ci / 4]
  __shared__ float sb[16];
  const int tid = threadIdx.x;
  if (tid < 64) {
    const uint2 u = *(const uint2*)(w + 4 * tid);
    sw[tid] = make_float4(bf2f((bf16_t)(u.x & 0xFFFF)), bf2f((bf16_t)(u.x >> 16)), bf2f((bf16_t)(u.y & 0xFFFF)),
                          bf2f((bf16_t)(u.y >> 16)));
  }
  if (tid < 16) sb[tid] = bias ? bias[tid] : 0.f;
  __syncthreads();
  for (long long v = blockIdx.x * (long long)blockDim.x + tid; v < N; v += (long long)gridDim.x * blockDim.x) {
    float2 res[MAX_MIX];
#pragma unroll
    for (int k = 0; k < MAX_MIX; ++k) res[k] = k < K ? c[(size_t)k * N + v] : make_float2(0.f, 0.f);
    const uint4* p = (const uint4*)(in + (size_t)v * 16);
    const uint4 u0 = p[0], u1 = p[1];
    const unsigned wd[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
    float x[16];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      x[2 * i] = __uint_as_float(wd[i] << 16);
      x[2 * i + 1] = __uint_as_float(wd[i] & 0xFFFF0000u);
    }
    float r[16];
#pragma unroll
    for (int co = 0; co < 16; ++co) {
      // weights read row by row (not hoisted into 256 registers), as in k_conv1_16
      int wi = co * 4;
      asm volatile("" : "+v"(wi));
      float acc = sb[co];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float4 wv = sw[wi + q];
        acc = fmaf(wv.x, x[4 * q], acc);
        acc = fmaf(wv.y, x[4 * q + 1], acc);
        acc = fmaf(wv.z, x[4 * q + 2], acc);
        acc = fmaf(wv.w, x[4 * q + 3], acc);
      }
      r[co] = relu ? fmaxf(acc, 0.f) : acc;
    }
#pragma unroll
    for (int k = 0; k < MAX_MIX; ++k) {
      if (k < K) {
        float2 z = make_float2(0.f, 0.f);
#pragma unroll
        for (int j = 0; j < MAX_MIX; ++j) {
          if (j < K) {
            z.x += M.m[k * MAX_MIX + j] * r[2 * j];
            z.y += M.m[k * MAX_MIX + j] * r[2 * j + 1];
          }
        }
        out[(size_t)k * N + v] = make_float2(fmaf(z.x, scale, res[k].x), fmaf(z.y, scale, res[k].y));
      }
    }
  }
}

// partial sums of |sum_k m0[k] c_k|^2 over voxels (the network's first channel pair)
__global__ void __launch_bounds__(256) k_coef_energy(const float2* __restrict__ c, long long N, int K, Mix M,
                                                     double* __restrict__ partial) {
  __shared__ double sh[4];
  double s = 0.0;
  for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < N; v += (long long)gridDim.x * blockDim.x) {
    float2 b = make_float2(0.f, 0.f);
#pragma unroll
    for (int k = 0; k < MAX_MIX; ++k) {
      if (k < K) {
        float2 z = c[(size_t)k * N + v];
        b.x += M.m[k] * z.x;
        b.y += M.m[k] * z.y;
      }
    }
    s += (double)b.x * b.x + (double)b.y * b.y;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// crop (cz, cy, cx) at origin (oz, oy, ox) of a (D, H, W, C) BF16 tensor -> FP32
__global__ void k_crop_bf16(const bf16_t* __restrict__ src, float* __restrict__ dst, int H, int W, int C, int oz,
                            int oy, int ox, int cz, int cy, int cx) {
  const long long total = (long long)cz * cy * cx * C;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x) {
    const int ch = (int)(e % C);
    long long v = e / C;
    const int x = (int)(v % cx), y = (int)((v / cx) % cy), z = (int)(v / ((long long)cx * cy));
    dst[e] = bf2f(src[((((size_t)(z + oz) * H) + y + oy) * W + x + ox) * C + ch]);
  }
}

}  // namespace

// =================================================================== network
struct UnetLayer {
  std::string name;
  int k, cin, cout;
  bool relu;
  size_t w_off, b_off;  // into the flat parameter vector (floats)
  size_t wb_off;        // into the BF16 packed weights (elements)
};

struct llr_unet {
  llr_ctx* ctx = nullptr;
  int cin = 16, base = 16, levels = 4;
  std::vector<UnetLayer> layers;
  size_t nparam = 0, nwb = 0;
  llr::DevBuf P, G, M, Vv, Wf, Wt, part, loss_d;
  long long step = 0;
  // activation / gradient workspace for one input size
  size_t dims[3] = {0, 0, 0};
  llr::DevBuf xb, xf, y, gy;
  std::vector<llr::DevBuf*> a, b, pool, cat, da, db;      // per level
  std::vector<llr::DevBuf*> ga, gb, gin, gcat, gskip, gda, gdb;
  ~llr_unet() {
    for (auto* v : {&a, &b, &pool, &cat, &da, &db, &ga, &gb, &gin, &gcat, &gskip, &gda, &gdb})
      for (llr::DevBuf* d : *v) delete d;
  }
};

namespace llr {

static std::vector<UnetLayer> unet_layers(int cin, int base, int levels) {
  std::vector<UnetLayer> L;
  std::vector<int> ch(levels);
  for (int l = 0; l < levels; ++l) ch[l] = base << l;
  int c = cin;
  for (int l = 0; l < levels; ++l) {
    L.push_back({"enc" + std::to_string(l) + "a", 3, c, ch[l], true, 0, 0, 0});
    L.push_back({"enc" + std::to_string(l) + "b", 3, ch[l], ch[l], true, 0, 0, 0});
    c = ch[l];
  }
  for (int l = levels - 2; l >= 0; --l) {
    L.push_back({"dec" + std::to_string(l) + "a", 3, c + ch[l], ch[l], true, 0, 0, 0});
    L.push_back({"dec" + std::to_string(l) + "b", 3, ch[l], ch[l], true, 0, 0, 0});
    c = ch[l];
  }
  L.push_back({"out", 1, c, cin, false, 0, 0, 0});
  size_t off = 0, wb = 0;
  for (UnetLayer& u : L) {
    size_t nw = (size_t)u.k * u.k * u.k * u.cout * u.cin;
    u.w_off = off;
    u.b_off = off + nw;
    u.wb_off = wb;
    off += nw + u.cout;
    wb += nw;
  }
  return L;
}

static const UnetLayer& layer(const llr_unet* U, const std::string& name) {
  for (const UnetLayer& u : U->layers)
    if (u.name == name) return u;
  return U->layers.back();
}

static llr_status pack_all(llr_unet* U) {
  for (const UnetLayer& u : U->layers) {
    int kk = u.k * u.k * u.k;
    size_t nw = (size_t)kk * u.cout * u.cin;
    hipLaunchKernelGGL(k_pack_weights, dim3(grid_for(nw, 256, 4096)), dim3(256), 0, U->ctx->stream,
                       U->P.as<float>() + u.w_off, U->Wf.as<bf16_t>() + u.wb_off, U->Wt.as<bf16_t>() + u.wb_off, kk,
                       u.cout, u.cin);
  }
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

static llr_status ensure_workspace(llr_unet* U, const size_t* dims) {
  if (U->dims[0] == dims[0] && U->dims[1] == dims[1] && U->dims[2] == dims[2]) return LLR_OK;
  const size_t f = (size_t)1 << (U->levels - 1);
  for (int d = 0; d < 3; ++d)
    if (dims[d] == 0 || dims[d] % f) return fail(LLR_ERR_ARG, "U-Net input sizes must be multiples of 2^(levels-1)");
  for (auto* v : {&U->a, &U->b, &U->pool, &U->cat, &U->da, &U->db, &U->ga, &U->gb, &U->gin, &U->gcat, &U->gskip,
                  &U->gda, &U->gdb}) {
    for (DevBuf* d : *v) delete d;
    v->assign(U->levels, nullptr);
    for (int l = 0; l < U->levels; ++l) (*v)[l] = new DevBuf();
  }
  const size_t V = dims[0] * dims[1] * dims[2];
  LLR_HIP(U->xb.alloc(V * U->cin * 2));
  LLR_HIP(U->xf.alloc(V * U->cin * 4));
  LLR_HIP(U->y.alloc(V * U->cin * 4));
  LLR_HIP(U->gy.alloc(V * U->cin * 4));
  for (int l = 0; l < U->levels; ++l) {
    const size_t Vl = V >> (3 * l);
    const size_t c = (size_t)U->base << l;
    const size_t cu = (size_t)U->base << (l + 1);
    LLR_HIP(U->a[l]->alloc(Vl * c * 2));
    LLR_HIP(U->b[l]->alloc(Vl * c * 2));
    LLR_HIP(U->ga[l]->alloc(Vl * c * 4));
    LLR_HIP(U->gb[l]->alloc(Vl * c * 4));
    if (l > 0) {
      const size_t cp = (size_t)U->base << (l - 1);
      LLR_HIP(U->pool[l]->alloc(Vl * cp * 2));
      LLR_HIP(U->gin[l]->alloc(Vl * cp * 4));
    }
    if (l < U->levels - 1) {
      LLR_HIP(U->cat[l]->alloc(Vl * (c + cu) * 2));
      LLR_HIP(U->gcat[l]->alloc(Vl * (c + cu) * 4));
      LLR_HIP(U->gskip[l]->alloc(Vl * c * 4));
      LLR_HIP(U->da[l]->alloc(Vl * c * 2));
      LLR_HIP(U->db[l]->alloc(Vl * c * 2));
      LLR_HIP(U->gda[l]->alloc(Vl * c * 4));
      LLR_HIP(U->gdb[l]->alloc(Vl * c * 4));
    }
  }
  for (int d = 0; d < 3; ++d) U->dims[d] = dims[d];
  return LLR_OK;
}

// dims are (nx, ny, nz) = (W, H, D)
template <typename TIn, typename TOut>
static llr_status conv(llr_unet* U, const UnetLayer& u, const size_t* dims, int lvl, const TIn* in, TOut* out,
                       bool backward_data, const bf16_t* mask, bool accumulate, const bf16_t* up = nullptr,
                       int cup = 0, const float* addend = nullptr) {
  ConvArgs a{};
  a.in = in;
  a.addend = addend;
  if (addend && (u.k != 1 || sizeof(TOut) != 4)) return fail(LLR_ERR_ARG, "addend: 1^3 FP32 output only");
  a.up = up;
  a.cup = up ? cup : 0;
  if (up && (u.k != 3 || backward_data || cup % 16)) return fail(LLR_ERR_ARG, "fused upsampling: 3^3 forward only");
  a.w = (backward_data ? U->Wt.as<bf16_t>() : U->Wf.as<bf16_t>()) + u.wb_off;
  a.bias = backward_data ? nullptr : U->P.as<float>() + u.b_off;
  a.mask = mask;
  a.out = out;
  a.W = (int)(dims[0] >> lvl);
  a.H = (int)(dims[1] >> lvl);
  a.D = (int)(dims[2] >> lvl);
  a.cin = backward_data ? u.cout : u.cin;
  a.cout = backward_data ? u.cin : u.cout;
  a.k = u.k;
  a.mirror = backward_data ? 1 : 0;
  a.relu = (!backward_data && u.relu) ? 1 : 0;
  a.accumulate = accumulate ? 1 : 0;
  if (a.cout > 16 * MAX_CO_TILES || a.cin % 16 || a.cout % 16) return fail(LLR_ERR_UNSUPPORTED, "U-Net channel count");
  const long long V = (long long)a.D * a.H * a.W;
  if ((double)V * std::max(a.cin, a.cout) >= 2147483647.0)
    return fail(LLR_ERR_UNSUPPORTED, "U-Net tensor larger than 2^31 values");
  if (a.k == 3) {
    // output channels in slices of at most 64 per workgroup (grid.y), so the
    // staged weights fit in LDS beside the input runs
    const int nt = a.cout / 16;
    const long long runs = (V + CV_M - 1) / CV_M;
    a.per_xcd = (int)((runs + 7) / 8);
    const unsigned gx = (unsigned)(8 * a.per_xcd);
    hipStream_t s = U->ctx->stream;
    const dim3 g4(gx, (unsigned)((nt + 3) / 4));
    // few runs (the deep levels): two channel tiles per workgroup over more
    // workgroups, at two resident per CU instead of one with four tiles
    const bool split = nt > 2 && runs < 2048;
    const dim3 g2(gx, (unsigned)((nt + 1) / 2));
    if (a.W % CV_M == 0) {
      if (nt <= 1)
        hipLaunchKernelGGL((k_conv3<TIn, TOut, 1, true>), dim3(gx), dim3(256), 0, s, a);
      else if (nt <= 2)
        hipLaunchKernelGGL((k_conv3<TIn, TOut, 2, true>), dim3(gx), dim3(256), 0, s, a);
      else
        hipLaunchKernelGGL((k_conv3<TIn, TOut, 4, true>), g4, dim3(256), 0, s, a);
    } else {
      if (nt <= 1)
        hipLaunchKernelGGL((k_conv3<TIn, TOut, 1, false>), dim3(gx), dim3(256), 0, s, a);
      else if (nt <= 2 || split)
        hipLaunchKernelGGL((k_conv3<TIn, TOut, 2, false>), g2, dim3(256), 0, s, a);
      else
        hipLaunchKernelGGL((k_conv3<TIn, TOut, 4, false>), g4, dim3(256), 0, s, a);
    }
  } else if constexpr (sizeof(TIn) == 2 && sizeof(TOut) == 4) {
    if (a.cin == 16 && a.cout == 16 && !a.mirror && !a.mask && !a.accumulate) {
      hipLaunchKernelGGL(k_conv1_16, dim3(grid_for((size_t)V, 256, 16384)), dim3(256), 0, U->ctx->stream,
                         (const bf16_t*)a.in, a.w, a.bias, a.addend, (float*)a.out, V, a.relu);
    } else {
      unsigned grid = (unsigned)((V + 63) / 64);
      hipLaunchKernelGGL((k_conv<TIn, TOut>), dim3(grid), dim3(256), 0, U->ctx->stream, a);
    }
  } else {
    unsigned grid = (unsigned)((V + 63) / 64);
    hipLaunchKernelGGL((k_conv<TIn, TOut>), dim3(grid), dim3(256), 0, U->ctx->stream, a);
  }
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

static llr_status wgrad(llr_unet* U, const UnetLayer& u, const size_t* dims, int lvl, const float* g,
                        const bf16_t* x) {
  const int W = (int)(dims[0] >> lvl), H = (int)(dims[1] >> lvl), D = (int)(dims[2] >> lvl);
  const long long V = (long long)D * H * W;
  const int tiles = (u.cout / 16) * (u.cin / 16);
  // about 2048 wavefronts per layer, at least 256 voxels each (fewer atomics)
  long long chunk = (V * tiles + 2047) / 2048;
  chunk = std::max<long long>(256, (chunk + 15) / 16 * 16);
  dim3 grid((unsigned)((V + chunk - 1) / chunk), (unsigned)tiles);
  if (u.k == 3)
    hipLaunchKernelGGL(k_conv_wgrad<27>, grid, dim3(64), 0, U->ctx->stream, g, x, U->G.as<float>() + u.w_off, D, H, W,
                       u.cin, u.cout, chunk);
  else if (u.k == 1)
    hipLaunchKernelGGL(k_conv_wgrad<1>, grid, dim3(64), 0, U->ctx->stream, g, x, U->G.as<float>() + u.w_off, D, H, W,
                       u.cin, u.cout, chunk);
  else
    return fail(LLR_ERR_UNSUPPORTED, "U-Net kernel size");
  hipLaunchKernelGGL(k_bias_grad, dim3(grid_for((size_t)V, 256 / u.cout, 1024)), dim3(256), 0, U->ctx->stream, g,
                     U->G.as<float>() + u.b_off, V, u.cout);
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

static unsigned ew_grid(size_t n) { return grid_for(n, 256, 16384); }

// forward of the packed input U->xb / U->xf; prediction (input + net) in U->y.
// keep_cat: store the decoder's [upsampled, skip] concatenations (the weight
// gradients read them); inference reads the two halves in the convolution.
static llr_status unet_forward(llr_unet* U, const size_t* dims, bool keep_cat = true, bool out_layer = true) {
  llr_status st;
  const int L = U->levels;
  const size_t W = dims[0], H = dims[1], D = dims[2];
  hipStream_t s = U->ctx->stream;
  for (int l = 0; l < L; ++l) {
    const bf16_t* in = l == 0 ? U->xb.as<bf16_t>() : U->pool[l]->as<bf16_t>();
    if ((st = conv(U, layer(U, "enc" + std::to_string(l) + "a"), dims, l, in, U->a[l]->as<bf16_t>(), false, nullptr,
                   false)))
      return st;
    if ((st = conv(U, layer(U, "enc" + std::to_string(l) + "b"), dims, l, U->a[l]->as<bf16_t>(),
                   U->b[l]->as<bf16_t>(), false, nullptr, false)))
      return st;
    if (l < L - 1) {
      const int c = U->base << l;
      size_t n = (W >> (l + 1)) * (H >> (l + 1)) * (D >> (l + 1)) * c;
      hipLaunchKernelGGL(k_pool_fwd, dim3(ew_grid(n)), dim3(256), 0, s, U->b[l]->as<bf16_t>(),
                         U->pool[l + 1]->as<bf16_t>(), (int)(D >> l), (int)(H >> l), (int)(W >> l), c);
    }
  }
  for (int l = L - 2; l >= 0; --l) {
    const bf16_t* coarse = (l == L - 2) ? U->b[L - 1]->as<bf16_t>() : U->db[l + 1]->as<bf16_t>();
    const int c = U->base << l, cu = U->base << (l + 1);
    if (keep_cat) {
      size_t n = (W >> l) * (H >> l) * (D >> l) * (size_t)(c + cu);
      hipLaunchKernelGGL(k_concat_fwd, dim3(ew_grid(n)), dim3(256), 0, s, coarse, U->b[l]->as<bf16_t>(),
                         U->cat[l]->as<bf16_t>(), (int)(D >> l), (int)(H >> l), (int)(W >> l), cu, c);
      if ((st = conv(U, layer(U, "dec" + std::to_string(l) + "a"), dims, l, U->cat[l]->as<bf16_t>(),
                     U->da[l]->as<bf16_t>(), false, nullptr, false)))
        return st;
    } else if ((st = conv(U, layer(U, "dec" + std::to_string(l) + "a"), dims, l, U->b[l]->as<bf16_t>(),
                          U->da[l]->as<bf16_t>(), false, nullptr, false, coarse, cu))) {
      return st;
    }
    if ((st = conv(U, layer(U, "dec" + std::to_string(l) + "b"), dims, l, U->da[l]->as<bf16_t>(),
                   U->db[l]->as<bf16_t>(), false, nullptr, false)))
      return st;
  }
  if (!out_layer) return LLR_OK;
  // prediction = input + 1^3 output convolution (residual), in one pass
  return conv(U, layer(U, "out"), dims, 0, U->db[0]->as<bf16_t>(), U->y.as<float>(), false, nullptr, false, nullptr, 0,
              U->xf.as<float>());
}

// gradients of sum(gy * prediction) into U->G (zeroed first)
static llr_status unet_backward(llr_unet* U, const size_t* dims) {
  llr_status st;
  const int L = U->levels;
  const size_t W = dims[0], H = dims[1], D = dims[2];
  hipStream_t s = U->ctx->stream;
  LLR_HIP(hipMemsetAsync(U->G.p, 0, U->nparam * 4, s));
  auto name = [](const char* p, int l, const char* q) { return std::string(p) + std::to_string(l) + q; };
  // output layer (input db[0], a ReLU output with one consumer)
  const UnetLayer& uo = layer(U, "out");
  if ((st = wgrad(U, uo, dims, 0, U->gy.as<float>(), U->db[0]->as<bf16_t>()))) return st;
  if ((st = conv(U, uo, dims, 0, U->gy.as<float>(), U->gdb[0]->as<float>(), true, U->db[0]->as<bf16_t>(), false)))
    return st;
  // decoder, finest level first; gdb[l] holds the pre-activation gradient of dec{l}b
  for (int l = 0; l < L - 1; ++l) {
    const UnetLayer& ub = layer(U, name("dec", l, "b"));
    const UnetLayer& ua = layer(U, name("dec", l, "a"));
    if ((st = wgrad(U, ub, dims, l, U->gdb[l]->as<float>(), U->da[l]->as<bf16_t>()))) return st;
    if ((st = conv(U, ub, dims, l, U->gdb[l]->as<float>(), U->gda[l]->as<float>(), true, U->da[l]->as<bf16_t>(),
                   false)))
      return st;
    if ((st = wgrad(U, ua, dims, l, U->gda[l]->as<float>(), U->cat[l]->as<bf16_t>()))) return st;
    if ((st = conv(U, ua, dims, l, U->gda[l]->as<float>(), U->gcat[l]->as<float>(), true, nullptr, false))) return st;
    const int c = U->base << l, cu = U->base << (l + 1);
    const bf16_t* coarse_act = (l == L - 2) ? U->b[L - 1]->as<bf16_t>() : U->db[l + 1]->as<bf16_t>();
    float* g_coarse = (l == L - 2) ? U->gb[L - 1]->as<float>() : U->gdb[l + 1]->as<float>();
    size_t n = ((W >> (l + 1)) * (H >> (l + 1)) * (D >> (l + 1))) * (size_t)cu + (W >> l) * (H >> l) * (D >> l) * c;
    hipLaunchKernelGGL(k_concat_bwd, dim3(ew_grid(n)), dim3(256), 0, s, U->gcat[l]->as<float>(), g_coarse, coarse_act,
                       U->gskip[l]->as<float>(), (int)(D >> l), (int)(H >> l), (int)(W >> l), cu, c);
  }
  // encoder, deepest level first; gb[l] holds the pre-activation gradient of enc{l}b
  for (int l = L - 1; l >= 0; --l) {
    if (l < L - 1) {
      // gb[l] = (gskip[l] + pool_bwd(gin[l+1])) * (b[l] > 0)
      const int c = U->base << l;
      size_t n = (W >> l) * (H >> l) * (D >> l) * (size_t)c;
      LLR_HIP(hipMemcpyAsync(U->gb[l]->p, U->gskip[l]->p, n * 4, hipMemcpyDeviceToDevice, s));
      hipLaunchKernelGGL(k_pool_bwd_add, dim3(ew_grid(n)), dim3(256), 0, s, U->gin[l + 1]->as<float>(),
                         U->gb[l]->as<float>(), U->b[l]->as<bf16_t>(), (int)(D >> l), (int)(H >> l), (int)(W >> l), c);
    }
    const UnetLayer& ub = layer(U, name("enc", l, "b"));
    const UnetLayer& ua = layer(U, name("enc", l, "a"));
    if ((st = wgrad(U, ub, dims, l, U->gb[l]->as<float>(), U->a[l]->as<bf16_t>()))) return st;
    if ((st = conv(U, ub, dims, l, U->gb[l]->as<float>(), U->ga[l]->as<float>(), true, U->a[l]->as<bf16_t>(), false)))
      return st;
    const bf16_t* in = l == 0 ? U->xb.as<bf16_t>() : U->pool[l]->as<bf16_t>();
    if ((st = wgrad(U, ua, dims, l, U->ga[l]->as<float>(), in))) return st;
    if (l > 0 &&
        (st = conv(U, ua, dims, l, U->ga[l]->as<float>(), U->gin[l]->as<float>(), true, nullptr, false)))
      return st;
  }
  return LLR_OK;
}

// pack the FP32 (V x cin) input held in U->xf into U->xb
static llr_status unet_set_input(llr_unet* U, const size_t* dims) {
  const size_t n = dims[0] * dims[1] * dims[2] * U->cin;
  hipLaunchKernelGGL(k_f32_to_bf16, dim3(ew_grid(n)), dim3(256), 0, U->ctx->stream, U->xf.as<float>(),
                     U->xb.as<bf16_t>(), (long long)n);
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

static Mix make_mix(const float* m, int K) {
  Mix M{};
  for (int j = 0; j < MAX_MIX; ++j)
    for (int k = 0; k < MAX_MIX; ++k)
      M.m[j * MAX_MIX + k] = (j < K && k < K) ? (m ? m[j * K + k] : (j == k ? 1.f : 0.f)) : 0.f;
  return M;
}

static unsigned vox_grid(size_t n) { return grid_for(n, 256, 16384); }

// U-Net warm start in subspace-coefficient space: K complex planes in `coef`
// (device) -> predicted planes in `out` (device; may alias coef). The network
// sees b = to_net * c / scale and its prediction p maps back to
// from_net * p * scale (NULL matrices are the identity).
llr_status unet_predict_device(llr_unet* U, const float2* coef, int K, const size_t* vol, float scale,
                               const float* to_net, const float* from_net, float2* out) {
  if (K < 1 || K > MAX_MIX || 2 * K > U->cin) return fail(LLR_ERR_ARG, "more subspace coefficients than network channels");
  llr_status st;
  if ((st = ensure_workspace(U, vol))) return st;
  const long long N = (long long)(vol[0] * vol[1] * vol[2]);
  const UnetLayer& uo = layer(U, "out");
  if (uo.k == 1 && uo.cin == 16 && uo.cout == 16) {
    // BF16 input channels only; the output layer, the unpacking and the
    // residual (the input coefficients) in one pass
    hipLaunchKernelGGL(k_pack_coef<float>, dim3(vox_grid((size_t)N)), dim3(256), 0, U->ctx->stream, coef,
                       (float*)nullptr, N, K, U->cin, make_mix(to_net, K), 1.0f / scale, U->xb.as<bf16_t>());
    if ((st = unet_forward(U, vol, false, false))) return st;
    hipLaunchKernelGGL(k_out_coef, dim3(vox_grid((size_t)N)), dim3(256), 0, U->ctx->stream,
                       U->db[0]->as<bf16_t>(), U->Wf.as<bf16_t>() + uo.wb_off, U->P.as<float>() + uo.b_off, coef,
                       out, N, K, make_mix(from_net, K), scale, uo.relu ? 1 : 0);
    LLR_HIP(hipGetLastError());
    return LLR_OK;
  }
  // the FP32 channels (the residual) and their BF16 copy (the network's input) in one pass
  hipLaunchKernelGGL(k_pack_coef<float>, dim3(vox_grid((size_t)N)), dim3(256), 0, U->ctx->stream, coef,
                     U->xf.as<float>(), N, K, U->cin, make_mix(to_net, K), 1.0f / scale, U->xb.as<bf16_t>());
  if ((st = unet_forward(U, vol, false))) return st;
  hipLaunchKernelGGL(k_unpack_coef, dim3(vox_grid((size_t)N)), dim3(256), 0, U->ctx->stream, U->y.as<float>(), out,
                     N, K, U->cin, make_mix(from_net, K), scale);
  LLR_HIP(hipGetLastError());
  return LLR_OK;
}

llr_ctx* unet_context(llr_unet* U) { return U->ctx; }

// sqrt(mean over voxels of |sum_k to_net[0][k] c_k|^2): the warm start's input scale
llr_status unet_coef_scale(llr_ctx* ctx, const float2* coef, int K, size_t N, const float* to_net, double* scale) {
  if (K < 1 || K > MAX_MIX) return fail(LLR_ERR_ARG, "subspace rank out of range");
  double* part = nullptr;
  const unsigned nb = 1024;
  llr_status st = red_scratch(ctx, nb, &part);
  if (st) return st;
  Mix M = make_mix(to_net, K);  // row 0 holds the first channel's weights
  hipLaunchKernelGGL(k_coef_energy, dim3(nb), dim3(256), 0, ctx->stream, coef, (long long)N, K, M, part);
  LLR_HIP(hipGetLastError());
  std::vector<double> h(nb);
  LLR_HIP(hipStreamSynchronize(ctx->stream));
  LLR_HIP(hipMemcpy(h.data(), part, nb * 8, hipMemcpyDeviceToHost));
  double t = 0;
  for (double v : h) t += v;
  *scale = std::sqrt(t / (double)N);
  return LLR_OK;
}

// One Adam step on the mean squared error: input staged in U->xf (FP32),
// target staged in U->gy (replaced by the loss gradient in place).
llr_status unet_adam_on_staged(llr_unet* U, const size_t* dims, float lr, double* loss) {
  llr_status st;
  const size_t n = dims[0] * dims[1] * dims[2] * U->cin;
  hipStream_t s = U->ctx->stream;
  if ((st = unet_set_input(U, dims)) || (st = unet_forward(U, dims))) return st;
  const unsigned nb = 1024;
  hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, s, U->y.as<float>(), U->gy.as<float>(), U->gy.as<float>(),
                     (long long)n, U->part.as<double>());
  LLR_HIP(hipGetLastError());
  if ((st = unet_backward(U, dims))) return st;
  U->step += 1;
  const double b1 = 0.9, b2 = 0.999, eps = 1e-8;
  const double c1 = 1.0 - std::pow(b1, (double)U->step), c2 = 1.0 - std::pow(b2, (double)U->step);
  // p -= lr * (m/c1) / (sqrt(v/c2) + eps) = lr_t * m / (sqrt(v) + eps_t)
  const float lr_t = (float)(lr * std::sqrt(c2) / c1), eps_t = (float)(eps * std::sqrt(c2));
  hipLaunchKernelGGL(k_adam, dim3(ew_grid(U->nparam)), dim3(256), 0, s, U->P.as<float>(), U->G.as<float>(),
                     U->M.as<float>(), U->Vv.as<float>(), (long long)U->nparam, lr_t, (float)b1, (float)b2, eps_t);
  LLR_HIP(hipGetLastError());
  if ((st = pack_all(U))) return st;
  if (loss) {
    std::vector<double> p(nb);
    LLR_HIP(hipStreamSynchronize(s));
    LLR_HIP(hipMemcpy(p.data(), U->part.p, nb * 8, hipMemcpyDeviceToHost));
    double t = 0;
    for (double v : p) t += v;
    *loss = t / (double)n;
  }
  return LLR_OK;
}

}  // namespace llr

using namespace llr;

extern "C" {

llr_status llr_unet_create(llr_ctx* ctx, int in_channels, int base_channels, int levels, unsigned long long seed,
                           llr_unet** out) {
  if (!ctx || !out) return fail(LLR_ERR_ARG, "null argument");
  if (in_channels <= 0 || in_channels % 16 || base_channels <= 0 || base_channels % 16 || levels < 1 || levels > 6)
    return fail(LLR_ERR_ARG, "channel counts must be multiples of 16 and levels in 1..6");
  {
    int widest = std::max(in_channels, base_channels << (levels - 1));
    if (levels > 1) widest = std::max(widest, (base_channels << (levels - 1)) + (base_channels << (levels - 2)));
    if (widest > 16 * MAX_CO_TILES) return fail(LLR_ERR_UNSUPPORTED, "network too wide for the convolution kernel");
  }
  LLR_HIP(hipSetDevice(ctx->device));
  llr_unet* U = new llr_unet();
  U->ctx = ctx;
  U->cin = in_channels;
  U->base = base_channels;
  U->levels = levels;
  U->layers = unet_layers(in_channels, base_channels, levels);
  const UnetLayer& last = U->layers.back();
  U->nparam = last.b_off + last.cout;
  U->nwb = last.wb_off + (size_t)last.k * last.k * last.k * last.cout * last.cin;
  // He-normal weights, zero output layer (the network starts as the identity map), zero biases
  std::vector<float> h(U->nparam, 0.f);
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  for (const UnetLayer& u : U->layers) {
    if (u.name == "out") continue;
    size_t nw = (size_t)u.k * u.k * u.k * u.cout * u.cin;
    double sd = std::sqrt(2.0 / ((double)u.k * u.k * u.k * u.cin));
    for (size_t i = 0; i < nw; ++i) h[u.w_off + i] = (float)(sd * nd(rng));
  }
  auto bail = [&](llr_status s) {
    delete U;
    return s;
  };
  if (U->P.alloc(U->nparam * 4) || U->G.alloc(U->nparam * 4) || U->M.alloc(U->nparam * 4) ||
      U->Vv.alloc(U->nparam * 4) || U->Wf.alloc(U->nwb * 2) || U->Wt.alloc(U->nwb * 2) || U->part.alloc(1024 * 8))
    return bail(fail(LLR_ERR_NOMEM, "U-Net parameter allocation failed"));
  if (hipMemcpy(U->P.p, h.data(), U->nparam * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(U->M.p, 0, U->nparam * 4) != hipSuccess || hipMemset(U->Vv.p, 0, U->nparam * 4) != hipSuccess)
    return bail(fail(LLR_ERR_HIP, "U-Net parameter upload failed"));
  llr_status st = pack_all(U);
  if (st) return bail(st);
  *out = U;
  return LLR_OK;
}

void llr_unet_destroy(llr_unet* U) {
  if (!U) return;
  (void)hipSetDevice(U->ctx->device);
  (void)hipStreamSynchronize(U->ctx->stream);
  delete U;
}

llr_status llr_unet_num_params(llr_unet* U, size_t* n) {
  if (!U || !n) return fail(LLR_ERR_ARG, "null argument");
  *n = U->nparam;
  return LLR_OK;
}

llr_status llr_unet_get_params(llr_unet* U, float* host, size_t n) {
  if (!U || !host || n != U->nparam) return fail(LLR_ERR_ARG, "parameter buffer must hold num_params floats");
  LLR_HIP(hipStreamSynchronize(U->ctx->stream));
  LLR_HIP(hipMemcpy(host, U->P.p, n * 4, hipMemcpyDeviceToHost));
  return LLR_OK;
}

llr_status llr_unet_set_params(llr_unet* U, const float* host, size_t n) {
  if (!U || !host || n != U->nparam) return fail(LLR_ERR_ARG, "parameter buffer must hold num_params floats");
  LLR_HIP(hipSetDevice(U->ctx->device));
  LLR_HIP(hipStreamSynchronize(U->ctx->stream));
  LLR_HIP(hipMemcpy(U->P.p, host, n * 4, hipMemcpyHostToDevice));
  LLR_HIP(hipMemset(U->M.p, 0, n * 4));
  LLR_HIP(hipMemset(U->Vv.p, 0, n * 4));
  U->step = 0;
  return pack_all(U);
}

llr_status llr_unet_forward(llr_unet* U, const float* x, float* y, const size_t dims[3], int on_device) {
  if (!U || !x || !y || !dims) return fail(LLR_ERR_ARG, "null argument");
  LLR_HIP(hipSetDevice(U->ctx->device));
  llr_status st;
  if ((st = ensure_workspace(U, dims))) return st;
  const size_t n = dims[0] * dims[1] * dims[2] * U->cin;
  hipStream_t s = U->ctx->stream;
  LLR_HIP(hipMemcpyAsync(U->xf.p, x, n * 4, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
  if ((st = unet_set_input(U, dims)) || (st = unet_forward(U, dims))) return st;
  LLR_HIP(hipMemcpyAsync(y, U->y.p, n * 4, on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
  LLR_HIP(hipStreamSynchronize(s));
  return LLR_OK;
}

llr_status llr_unet_gradients(llr_unet* U, const float* x, const float* gy, const size_t dims[3], float* grads) {
  if (!U || !x || !gy || !dims || !grads) return fail(LLR_ERR_ARG, "null argument");
  LLR_HIP(hipSetDevice(U->ctx->device));
  llr_status st;
  if ((st = ensure_workspace(U, dims))) return st;
  const size_t n = dims[0] * dims[1] * dims[2] * U->cin;
  hipStream_t s = U->ctx->stream;
  LLR_HIP(hipMemcpyAsync(U->xf.p, x, n * 4, hipMemcpyHostToDevice, s));
  LLR_HIP(hipMemcpyAsync(U->gy.p, gy, n * 4, hipMemcpyHostToDevice, s));
  if ((st = unet_set_input(U, dims)) || (st = unet_forward(U, dims)) || (st = unet_backward(U, dims))) return st;
  LLR_HIP(hipStreamSynchronize(s));
  LLR_HIP(hipMemcpy(grads, U->G.p, U->nparam * 4, hipMemcpyDeviceToHost));
  return LLR_OK;
}

llr_status llr_unet_train_step(llr_unet* U, const float* x, const float* target, const size_t dims[3], float lr,
                               int on_device, double* loss) {
  if (!U || !x || !target || !dims) return fail(LLR_ERR_ARG, "null argument");
  LLR_HIP(hipSetDevice(U->ctx->device));
  llr_status st;
  if ((st = ensure_workspace(U, dims))) return st;
  const size_t n = dims[0] * dims[1] * dims[2] * U->cin;
  hipStream_t s = U->ctx->stream;
  const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  LLR_HIP(hipMemcpyAsync(U->xf.p, x, n * 4, kind, s));
  LLR_HIP(hipMemcpyAsync(U->gy.p, target, n * 4, kind, s));
  return unet_adam_on_staged(U, dims, lr, loss);
}

llr_status llr_unet_get_activation(llr_unet* U, int kind, int level, float* host, size_t n) {
  if (!U || !host || level < 0 || level >= U->levels) return fail(LLR_ERR_ARG, "bad activation request");
  std::vector<DevBuf*>* t[6] = {&U->a, &U->b, &U->pool, &U->cat, &U->da, &U->db};
  if (kind < 0 || kind > 5 || U->dims[0] == 0) return fail(LLR_ERR_ARG, "no such activation");
  DevBuf* d = (*t[kind])[level];
  if (!d || !d->p || d->bytes != n * 2) return fail(LLR_ERR_ARG, "activation size mismatch");
  std::vector<unsigned short> h(n);
  LLR_HIP(hipStreamSynchronize(U->ctx->stream));
  LLR_HIP(hipMemcpy(h.data(), d->p, n * 2, hipMemcpyDeviceToHost));
  for (size_t i = 0; i < n; ++i) {
    unsigned u = (unsigned)h[i] << 16;
    std::memcpy(host + i, &u, 4);
  }
  return LLR_OK;
}

llr_status llr_unet_predict_coefficients(llr_unet* U, const llr_cf32* d_coef, size_t K, const size_t vol[3],
                                         float scale, const float* to_net, const float* from_net, llr_cf32* d_out) {
  if (!U || !d_coef || !vol || !d_out) return fail(LLR_ERR_ARG, "null argument");
  if (!(scale > 0.f)) return fail(LLR_ERR_ARG, "scale must be positive");
  LLR_HIP(hipSetDevice(U->ctx->device));
  llr_status st = unet_predict_device(U, (const float2*)d_coef, (int)K, vol, scale, to_net, from_net, (float2*)d_out);
  if (st) return st;
  LLR_HIP(hipStreamSynchronize(U->ctx->stream));
  return LLR_OK;
}

llr_status llr_unet_pack_coefficients(llr_ctx* ctx, const llr_cf32* d_coef, size_t K, const size_t vol[3],
                                      int channels, float scale, const float* to_net, int bf16_out, void* d_out) {
  if (!ctx || !d_coef || !vol || !d_out) return fail(LLR_ERR_ARG, "null argument");
  if (K < 1 || K > (size_t)MAX_MIX || channels < 2 * MAX_MIX || channels % 16 || !(scale > 0.f))
    return fail(LLR_ERR_ARG, "bad subspace rank, channel count or scale");
  LLR_HIP(hipSetDevice(ctx->device));
  const long long N = (long long)(vol[0] * vol[1] * vol[2]);
  Mix M = make_mix(to_net, (int)K);
  if (bf16_out)
    hipLaunchKernelGGL(k_pack_coef<bf16_t>, dim3(vox_grid((size_t)N)), dim3(256), 0, ctx->stream,
                       (const float2*)d_coef, (bf16_t*)d_out, N, (int)K, channels, M, 1.0f / scale);
  else
    hipLaunchKernelGGL(k_pack_coef<float>, dim3(vox_grid((size_t)N)), dim3(256), 0, ctx->stream,
                       (const float2*)d_coef, (float*)d_out, N, (int)K, channels, M, 1.0f / scale);
  LLR_HIP(hipGetLastError());
  LLR_HIP(hipStreamSynchronize(ctx->stream));
  return LLR_OK;
}

llr_status llr_unet_coefficient_scale(llr_ctx* ctx, const llr_cf32* d_coef, size_t K, const size_t vol[3],
                                      const float* to_net, double* scale) {
  if (!ctx || !d_coef || !vol || !scale) return fail(LLR_ERR_ARG, "null argument");
  LLR_HIP(hipSetDevice(ctx->device));
  return unet_coef_scale(ctx, (const float2*)d_coef, (int)K, vol[0] * vol[1] * vol[2], to_net, scale);
}

llr_status llr_unet_train_step_crop(llr_unet* U, const void* x_full, const void* t_full, const size_t full_dims[3],
                                    const size_t origin[3], const size_t crop_dims[3], float lr, double* loss) {
  if (!U || !x_full || !t_full || !full_dims || !origin || !crop_dims) return fail(LLR_ERR_ARG, "null argument");
  for (int d = 0; d < 3; ++d)
    if (crop_dims[d] == 0 || origin[d] + crop_dims[d] > full_dims[d])
      return fail(LLR_ERR_ARG, "crop outside the stored tensor");
  LLR_HIP(hipSetDevice(U->ctx->device));
  llr_status st;
  if ((st = ensure_workspace(U, crop_dims))) return st;
  const size_t n = crop_dims[0] * crop_dims[1] * crop_dims[2] * U->cin;
  hipStream_t s = U->ctx->stream;
  const int W = (int)full_dims[0], H = (int)full_dims[1];
  const int ox = (int)origin[0], oy = (int)origin[1], oz = (int)origin[2];
  const int cx = (int)crop_dims[0], cy = (int)crop_dims[1], cz = (int)crop_dims[2];
  hipLaunchKernelGGL(k_crop_bf16, dim3(ew_grid(n)), dim3(256), 0, s, (const bf16_t*)x_full, U->xf.as<float>(), H, W,
                     U->cin, oz, oy, ox, cz, cy, cx);
  hipLaunchKernelGGL(k_crop_bf16, dim3(ew_grid(n)), dim3(256), 0, s, (const bf16_t*)t_full, U->gy.as<float>(), H, W,
                     U->cin, oz, oy, ox, cz, cy, cx);
  LLR_HIP(hipGetLastError());
  return unet_adam_on_staged(U, crop_dims, lr, loss);
}

}  // extern "C"
