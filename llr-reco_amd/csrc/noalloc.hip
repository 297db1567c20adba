// Allocation-free form of the device signal model (llr_signal_model_device):
// the twiddle tables, the per-volume phase-ramp tables and the corrections
// live in one file-scope device array instead of buffers allocated per call
// or cached with hipMalloc, so callers that cannot allocate (code under graph
// capture, tight serving loops) can run the forward model and its adjoint on
// their own device buffers.
//
// Each GPU holds its own copy of `scratch` for the life of the process. Every
// call on a GPU uses that GPU's copy: calls that overlap on one GPU (two host
// threads, or two contexts' streams) race on the tables.
#include <cmath>
#include <mutex>
#include <string>

#include "common.h"
#include "internal.h"

// 2 MiB: twiddles of the three axes, then the k-space and image ramp tables of
// every volume (n_volumes * (nx + ny + nz) values each), then the corrections
constexpr size_t kScratchElems = 262144;  // complex values
static __device__ __align__(256) float2 scratch[kScratchElems];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
float2* g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device

// The device address of this GPU's `scratch` (looked up once per device with
// that device current; the host-side name is not a device address).
llr_status scratch_on(int device, float2** out) {
  if (device < 0 || device >= kMaxDevices) return llr::fail(LLR_ERR_UNSUPPORTED, "device number out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return llr::hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return llr::fail(LLR_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
    g_scratch[device] = (float2*)p;
  }
  *out = g_scratch[device];
  return LLR_OK;
}

struct TableArgs {
  float2* s;         // scratch
  int n[3];          // axis lengths
  int nt;            // n[0] + n[1] + n[2]
  int nv;            // volumes
  long long pk0;     // first k-space ramp table (after the twiddles)
  long long pi0;     // first image ramp table
  const float* ksp;  // nv x 3 corrections in scratch, or null (zero)
  const float* img;
};

// exp(-2 pi i m / n) twiddles, and exp(+2 pi i frac((i - n/2) s / n)) ramps with
// s = c + n/2 (signal_model.rs:110-117, cufft.rs:320-326), the turn count
// reduced in FP64 before sincospi. Axis a's tables start at the prefix sum of
// the lengths before it; volume v's ramp at v * n[a] inside its axis block.
__global__ void k_noalloc_tables(TableArgs t) {
  const long long total = (long long)t.nt + 2LL * t.nv * t.nt;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    double sn, cs;
    if (e < t.nt) {
      const int a = e < t.n[0] ? 0 : e < t.n[0] + t.n[1] ? 1 : 2;
      const int m = (int)e - (a > 0 ? t.n[0] : 0) - (a > 1 ? t.n[1] : 0);
      sincospi(-2.0 * (double)m / (double)t.n[a], &sn, &cs);
      t.s[e] = make_float2((float)cs, (float)sn);
      continue;
    }
    const long long r = e - t.nt;
    const bool img = r >= (long long)t.nv * t.nt;
    const long long q = img ? r - (long long)t.nv * t.nt : r;
    // axis block a holds nv * n[a] values
    const long long b0 = (long long)t.nv * t.n[0], b1 = b0 + (long long)t.nv * t.n[1];
    const int a = q < b0 ? 0 : q < b1 ? 1 : 2;
    const long long w = q - (a == 0 ? 0 : a == 1 ? b0 : b1);
    const int n = t.n[a];
    const int v = (int)(w / n), i = (int)(w - (long long)v * n);
    const float* c = img ? t.img : t.ksp;
    const double s = (double)(c ? c[3 * v + a] : 0.0f) + 0.5 * (double)n;
    double tr = ((double)i - 0.5 * (double)n) * s / (double)n;
    tr -= floor(tr);
    sincospi(2.0 * tr, &sn, &cs);
    t.s[e] = make_float2((float)cs, (float)sn);
  }
}

}  // namespace

using namespace llr;

extern "C" {

llr_status llr_signal_model_device_noalloc(llr_ctx* ctx, llr_cf32* x, size_t len, const size_t image_size[3],
                                           size_t batch_size, const float* lin_ksp_corr, const float* lin_img_corr,
                                           int direction) {
  if (!ctx || !x || !image_size) return fail(LLR_ERR_ARG, "null argument");
  const size_t N = image_size[0] * image_size[1] * image_size[2];
  if (len != N * batch_size) return fail(LLR_ERR_ARG, "data layout and data length inconsistency");
  if (direction != LLR_FORWARD && direction != LLR_INVERSE) return fail(LLR_ERR_ARG, "bad direction");
  for (int a = 0; a < 3; ++a)
    if (image_size[a] < 2 || !fft_size_supported(image_size[a]))
      return fail(LLR_ERR_UNSUPPORTED, "volume dimension " + std::to_string(image_size[a]) +
                                           " is not a supported FFT length (2..512, radix 2/3)");
  if (batch_size == 0) return LLR_OK;
  const size_t nt = image_size[0] + image_size[1] + image_size[2];
  // twiddles, two ramp tables per volume, then 2 x nv x 3 floats of corrections
  const size_t corr_elems = 3 * batch_size;  // complex slots holding 6 floats per volume
  const size_t need = nt + 2 * batch_size * nt + corr_elems;
  if (batch_size > kScratchElems || need > kScratchElems)
    return fail(LLR_ERR_UNSUPPORTED, "allocation-free signal model: " + std::to_string(batch_size) + " volumes of " +
                                         std::to_string(image_size[0]) + "x" + std::to_string(image_size[1]) + "x" +
                                         std::to_string(image_size[2]) + " need " + std::to_string(need) +
                                         " table entries, the scratch array holds " + std::to_string(kScratchElems) +
                                         " (volumes * (2 (nx+ny+nz) + 3) + nx+ny+nz at most)");
  LLR_HIP(hipSetDevice(ctx->device));
  float2* S = nullptr;
  llr_status st = scratch_on(ctx->device, &S);
  if (st) return st;
  TableArgs t{};
  t.s = S;
  for (int a = 0; a < 3; ++a) t.n[a] = (int)image_size[a];
  t.nt = (int)nt;
  t.nv = (int)batch_size;
  t.pk0 = (long long)nt;
  t.pi0 = (long long)(nt + batch_size * nt);
  float* corr = (float*)(S + nt + 2 * batch_size * nt);
  if (lin_ksp_corr) {
    LLR_HIP(hipMemcpyAsync(corr, lin_ksp_corr, 3 * batch_size * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    t.ksp = corr;
  }
  if (lin_img_corr) {
    LLR_HIP(hipMemcpyAsync(corr + 3 * batch_size, lin_img_corr, 3 * batch_size * sizeof(float),
                           hipMemcpyHostToDevice, ctx->stream));
    t.img = corr + 3 * batch_size;
  }
  const size_t entries = nt + 2 * batch_size * nt;
  hipLaunchKernelGGL(k_noalloc_tables, dim3(grid_for(entries, 256, 1024u)), dim3(256), 0, ctx->stream, t);
  LLR_HIP(hipGetLastError());
  // the three ramped passes of signal_model.rs:34-79 (forward: k-space ramp in,
  // image ramp conjugated out; adjoint: the reverse), tables from the scratch
  size_t off_n = 0, off_v = 0;
  for (int a = 0; a < 3; ++a) {
    const float2* tw = S + off_n;
    const float2* pk = S + t.pk0 + off_v;
    const float2* pim = S + t.pi0 + off_v;
    const float sc = (float)(1.0 / std::sqrt((double)image_size[a]));
    if (direction == LLR_FORWARD)
      st = fft_axis_pass(ctx, x, x, image_size, batch_size, a, false, pk, false, pim, true, sc, false, tw);
    else
      st = fft_axis_pass(ctx, x, x, image_size, batch_size, a, true, pim, false, pk, true, sc, false, tw);
    if (st) return st;
    off_n += image_size[a];
    off_v += batch_size * image_size[a];
  }
  LLR_HIP(hipStreamSynchronize(ctx->stream));
  return LLR_OK;
}

}  // extern "C"
