// gsk_reduce_eta_noalloc: the mean-eta and sum-of-squares reductions of gsk_reduce on order
// parameters the caller already holds in device memory, for callers that cannot allocate
// (tight serving loops, harnesses that forbid hipMalloc). It allocates nothing: its per-block
// partial sums live in the file-scope device array `scratch`, which every call on a GPU shares.
#include "gsk_internal.h"

#include <mutex>

namespace {

// partial sums of kNaBlocks blocks x (2 + kNaMaxVar) slots, then the finished sums
constexpr int kNaBlocks = 512;  // the block count of gsk_reduce's real-space pass: same sums, same order
constexpr int kNaMaxVar = 8;
constexpr int kNaSlots = 2 + kNaMaxVar;
constexpr size_t kNaDoubles = (size_t)kNaBlocks * kNaSlots + 16;  // 41,088 bytes

}  // namespace

static __device__ __align__(256) double scratch[kNaDoubles];

namespace {

constexpr int kNaMaxDev = 64;
std::mutex g_na_mu;
double* g_na_ptr[kNaMaxDev] = {};  // device address of `scratch` on each GPU

// Device address of this GPU's copy of `scratch` (the current device must be `dev`); looked up
// once per device. The host-side name `scratch` is not a device address and is never passed.
int na_scratch(int dev, double** out) {
  if (dev < 0 || dev >= kNaMaxDev) return fail(GSK_ERR_ARG, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (!g_na_ptr[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_na_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_na_ptr[dev];
  return GSK_OK;
}

}  // namespace

extern "C" int gsk_reduce_eta_noalloc(const float* d_eta, int nvar, long long nvox, int what, double* out,
                                      int device) {
  if (!d_eta || !out) return fail(GSK_ERR_ARG, "null argument");
  if (what != GSK_RED_MEAN_ETA && what != GSK_RED_SUM_ETA2)
    return fail(GSK_ERR_ARG, "gsk_reduce_eta_noalloc computes GSK_RED_MEAN_ETA or GSK_RED_SUM_ETA2");
  if (nvar < 1 || nvar > kNaMaxVar)
    return fail(GSK_ERR_ARG, "gsk_reduce_eta_noalloc: nvar must be 1.." + std::to_string(kNaMaxVar) +
                                 " (its scratch holds " + std::to_string(kNaSlots) + " sums per block)");
  if (nvox <= 0) return fail(GSK_ERR_ARG, "gsk_reduce_eta_noalloc: nvox must be positive");
  CHK(ensure_device(device));
  double* sc = nullptr;
  CHK(na_scratch(device, &sc));
  // every block writes all of its 2 + nvar partial slots and the finishing block writes every
  // output slot, so nothing in the scratch has to start at zero
  double* fin = sc + (size_t)kNaBlocks * kNaSlots;
  hipLaunchKernelGGL(r_mart_real<float>, dim3(kNaBlocks), dim3(256), 0, 0, d_eta, (long)nvox, nvar, 0.0, 0.0, 0.0, sc);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(r_finish<>, dim3(1), dim3(1024), 0, 0, sc, kNaBlocks, 2 + nvar, fin);
  HIPCHK(hipGetLastError());
  double h[kNaSlots];
  HIPCHK(hipMemcpy(h, fin, (2 + nvar) * sizeof(double), hipMemcpyDeviceToHost));
  if (what == GSK_RED_SUM_ETA2) {
    out[0] = h[1];
  } else {
    for (int v = 0; v < nvar; ++v) out[v] = h[2 + v] / (double)nvox;
  }
  return GSK_OK;
}
