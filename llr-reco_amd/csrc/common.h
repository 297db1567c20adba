// Shared device helpers and host-side context for the LLR HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/llr_hip.h"

// ------------------------------------------------------------ complex math
// Interleaved complex64 as float2; textbook formulas (no inf/nan recovery).
__host__ __device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__host__ __device__ inline float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// conj(a) * b
__host__ __device__ inline float2 cconjmul(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__host__ __device__ inline float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__host__ __device__ inline float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__host__ __device__ inline float cnorm(float2 a) { return a.x * a.x + a.y * a.y; }

// ------------------------------------------------------- complex storage
// Large buffers hold complex values either as complex64 (float2) or as two
// IEEE half-precision numbers in one 32-bit word (low half: real part).
// Kernels templated on the storage load into FP32 registers (exactly) and
// round once, to nearest even, when storing; all arithmetic is FP32. Data in
// half storage is kept at a scale where every value is below 2^14 in
// magnitude (the caller's job), far inside the format's range.
// Prefetching kernels hold the raw words (T, from zero() or a plain load) and
// widen them with cvt() only where the values are used: a widening next to
// the load would make the wave wait for the load at once.
// ldnt / rawnt / stnt: non-temporal loads and stores (streamed data read or
// written once per pass, far larger than the caches).
typedef unsigned int ch16;
typedef _Float16 half2v __attribute__((ext_vector_type(2)));
typedef float f32x2v __attribute__((ext_vector_type(2)));

struct StoreF32 {
  using T = float2;
  static constexpr int bytes = 8;
  __device__ static inline float2 ld(const float2* p) { return *p; }
  __device__ static inline void st(float2* p, float2 v) { *p = v; }
  __device__ static inline float2 cvt(float2 v) { return v; }
  __device__ static inline float2 zero() { return make_float2(0.f, 0.f); }
  __device__ static inline float2 rawnt(const float2* p) {
    const f32x2v v = __builtin_nontemporal_load((const f32x2v*)p);
    return make_float2(v.x, v.y);
  }
  __device__ static inline float2 ldnt(const float2* p) { return rawnt(p); }
  __device__ static inline void stnt(float2* p, float2 v) {
    const f32x2v w = {v.x, v.y};
    __builtin_nontemporal_store(w, (f32x2v*)p);
  }
};

struct StoreF16 {
  using T = ch16;
  static constexpr int bytes = 4;
  __device__ static inline float2 unpack(ch16 u) {
    const half2v h = __builtin_bit_cast(half2v, u);
    return make_float2((float)h.x, (float)h.y);
  }
  __device__ static inline ch16 pack(float2 v) {
    const half2v h = {(_Float16)v.x, (_Float16)v.y};
    return __builtin_bit_cast(ch16, h);
  }
  __device__ static inline float2 ld(const ch16* p) { return unpack(*p); }
  __device__ static inline void st(ch16* p, float2 v) { *p = pack(v); }
  __device__ static inline float2 cvt(ch16 u) { return unpack(u); }
  __device__ static inline ch16 zero() { return 0u; }
  __device__ static inline ch16 rawnt(const ch16* p) { return __builtin_nontemporal_load(p); }
  __device__ static inline float2 ldnt(const ch16* p) { return unpack(rawnt(p)); }
  __device__ static inline void stnt(ch16* p, float2 v) { __builtin_nontemporal_store(pack(v), p); }
};

// ------------------------------------------------------------- host context
struct llr_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // cached FFT twiddles exp(-2*pi*i*m/n), m in [0, n), keyed by n
  std::vector<std::pair<int, float2*>> twiddles;
  // per-call scratch for reductions
  double* d_red = nullptr;
  size_t red_cap = 0;
  // launch-shape overrides chosen by the autotuner (0 / -1: built-in default)
  int fft_tpw[3] = {0, 0, 0};  // FFT tiles per workgroup, per axis (z: the fused projection pass)
  int llr_stage = -1;          // LLR block kernel: 1 stage the block in LDS, 0 read it from HBM twice
  int lift_vpw = 0;           // gather/scatter slab kernels: volumes per workgroup (0: built-in default)
  int proj_tpw = 0;           // half-storage projection: 128-voxel tiles per workgroup (0: built-in default)
  int wav_tpw = 0;            // wavelet xy synthesis: x tiles per workgroup (0: built-in default)
  // per-kernel timing inside multi-kernel stages: while a reconstruction's
  // stage timing is on these point at its stage tables (null otherwise)
  double* kt_ms = nullptr;
  size_t* kt_n = nullptr;
  hipEvent_t kt_ev[2] = {nullptr, nullptr};
};

namespace llr {

void set_error(const std::string& msg);
llr_status fail(llr_status st, const std::string& msg);
llr_status hip_fail(hipError_t e, const char* what);

#define LLR_HIP(call)                                         \
  do {                                                        \
    hipError_t e_ = (call);                                   \
    if (e_ != hipSuccess) return ::llr::hip_fail(e_, #call);  \
  } while (0)

#define LLR_CHECK(cond, msg)                                       \
  do {                                                             \
    if (!(cond)) return ::llr::fail(LLR_ERR_ARG, (msg));           \
  } while (0)

// RAII device buffer used by the host-pointer entry points.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t b) {
    if (p) {
      (void)hipFree(p);
      p = nullptr;
    }
    bytes = b;
    return hipMalloc(&p, b ? b : 1);
  }
  template <class T>
  T* as() const {
    return (T*)p;
  }
};

// Twiddle table of length n (exp(-2*pi*i*m/n)) computed in long double on the
// host; cached per context.
llr_status get_twiddles(llr_ctx* ctx, int n, const float2** out);

// Launch-shape helper: blocks for `total` work items, capped for grid-stride loops.
inline unsigned grid_for(size_t total, unsigned block, unsigned cap = 65535u * 8u) {
  size_t g = (total + block - 1) / block;
  if (g == 0) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

// Reduction scratch of `n` doubles on the context's device.
llr_status red_scratch(llr_ctx* ctx, size_t n, double** out);

// Times the launches made in its scope into stage `id` of the context's
// timing tables when they are set (HIP events on the context's stream; waits
// for the launches, so only for timing runs).
struct KernelTimer {
  llr_ctx* c;
  int id;
  KernelTimer(llr_ctx* c_, int id_) : c(c_), id(id_) {
    if (c->kt_ms) (void)hipEventRecord(c->kt_ev[0], c->stream);
  }
  ~KernelTimer() {
    if (!c->kt_ms) return;
    (void)hipEventRecord(c->kt_ev[1], c->stream);
    (void)hipEventSynchronize(c->kt_ev[1]);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->kt_ev[0], c->kt_ev[1]);
    c->kt_ms[id] += ms;
    c->kt_n[id] += 1;
  }
};
// stage ids of the per-kernel timers (see llr_recon_stage_times)
constexpr int kStageWavIn = 11, kStageWavMid = 12, kStageWavOut = 13, kStageWavDeep = 14;

}  // namespace llr
