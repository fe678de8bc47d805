// Timing comparison of the library's 3D R2C / C2R transforms with rocFFT on the
// same data (rocFFT is only the comparator; the step never calls it).
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <cmath>
#include <string>
#include <vector>

#include "gsk.h"

extern "C" int gsk_fft_prepare_(int n, int count, int device, void** sim, float** real, float2** kbuf,
                                hipStream_t* stream);
extern "C" int gsk_fft_run_(void* sim, int dir);
extern "C" int gsk_fft_release_(void* sim);
extern "C" int gsk_fft_spectrum_(void* sim, float* host);
extern "C" int gsk_fail_(int code, const char* msg);

namespace {
struct Ev {
  hipEvent_t a{}, b{};
  Ev() { (void)hipEventCreate(&a); (void)hipEventCreate(&b); }
  ~Ev() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
};
}  // namespace

extern "C" int gsk_fft_compare(int n, int count, int reps, double* out, int device) {
  // out[0] hand R2C ms, out[1] rocFFT R2C ms, out[2] hand C2R ms, out[3] rocFFT C2R ms (per call of
  // `count` fields), out[4] max |hand - rocFFT| / max |rocFFT| of the R2C spectra.
  if (!out || reps < 1 || count < 1) return gsk_fail_(GSK_ERR_ARG, "bad argument");
  void* sim = nullptr;
  float* real = nullptr;
  float2* kbuf = nullptr;
  hipStream_t stream = nullptr;
  int rc = gsk_fft_prepare_(n, count, device, &sim, &real, &kbuf, &stream);
  if (rc != GSK_OK) return rc;
  const size_t N3 = (size_t)n * n * n, H = (size_t)n * n * (n / 2 + 1);
  float* rin = nullptr;
  float2* rout = nullptr;
  void* work = nullptr;
  rocfft_plan pf = nullptr, pb = nullptr;
  rocfft_execution_info info = nullptr;
  std::string err;
  auto cleanup = [&]() {
    if (pf) rocfft_plan_destroy(pf);
    if (pb) rocfft_plan_destroy(pb);
    if (info) rocfft_execution_info_destroy(info);
    if (rin) (void)hipFree(rin);
    if (rout) (void)hipFree(rout);
    if (work) (void)hipFree(work);
    gsk_fft_release_(sim);
  };
  do {
    if (rocfft_setup() != rocfft_status_success) { err = "rocfft_setup failed"; break; }
    if (hipMalloc((void**)&rin, count * N3 * sizeof(float)) != hipSuccess) { err = "hipMalloc"; break; }
    if (hipMalloc((void**)&rout, count * H * sizeof(float2)) != hipSuccess) { err = "hipMalloc"; break; }
    std::vector<float> h(count * N3);
    for (size_t i = 0; i < h.size(); ++i) h[i] = (float)std::sin(0.001 * (double)i) + 0.25f * (float)((i * 2654435761u) % 1000) / 1000.f;
    if (hipMemcpy(rin, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) { err = "copy"; break; }
    if (hipMemcpy(real, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) { err = "copy"; break; }
    // rocFFT lengths are given fastest dimension first
    const size_t len[3] = {(size_t)n, (size_t)n, (size_t)n};
    if (rocfft_plan_create(&pf, rocfft_placement_notinplace, rocfft_transform_type_real_forward,
                           rocfft_precision_single, 3, len, count, nullptr) != rocfft_status_success ||
        rocfft_plan_create(&pb, rocfft_placement_notinplace, rocfft_transform_type_real_inverse,
                           rocfft_precision_single, 3, len, count, nullptr) != rocfft_status_success) {
      err = "rocfft_plan_create failed";
      break;
    }
    size_t w1 = 0, w2 = 0;
    rocfft_plan_get_work_buffer_size(pf, &w1);
    rocfft_plan_get_work_buffer_size(pb, &w2);
    const size_t ws = w1 > w2 ? w1 : w2;
    rocfft_execution_info_create(&info);
    rocfft_execution_info_set_stream(info, stream);
    if (ws) {
      if (hipMalloc(&work, ws) != hipSuccess) { err = "hipMalloc work"; break; }
      rocfft_execution_info_set_work_buffer(info, work, ws);
    }
    // correctness: forward spectra of the same data
    void* ib[1] = {rin};
    void* ob[1] = {rout};
    if (rocfft_execute(pf, ib, ob, info) != rocfft_status_success) { err = "rocfft_execute"; break; }
    if ((rc = gsk_fft_run_(sim, +1)) != GSK_OK) { cleanup(); return rc; }
    if (hipStreamSynchronize(stream) != hipSuccess) { err = "sync"; break; }
    {
      std::vector<float2> a(count * H), bk(count * H);
      if (hipMemcpy(a.data(), rout, a.size() * sizeof(float2), hipMemcpyDeviceToHost) != hipSuccess ||
          gsk_fft_spectrum_(sim, (float*)bk.data()) != GSK_OK) {
        err = "copy back";
        break;
      }
      double md = 0, mx = 0;
      for (size_t i = 0; i < a.size(); ++i) {
        md = std::fmax(md, std::hypot(a[i].x - bk[i].x, a[i].y - bk[i].y));
        mx = std::fmax(mx, std::hypot(a[i].x, a[i].y));
      }
      out[4] = md / mx;
    }
    Ev ev;
    float t = 0;
    // rocFFT forward / inverse
    (void)hipEventRecord(ev.a, stream);
    for (int i = 0; i < reps; ++i) rocfft_execute(pf, ib, ob, info);
    (void)hipEventRecord(ev.b, stream);
    (void)hipEventSynchronize(ev.b);
    (void)hipEventElapsedTime(&t, ev.a, ev.b);
    out[1] = t / reps;
    (void)hipEventRecord(ev.a, stream);
    for (int i = 0; i < reps; ++i) rocfft_execute(pb, ob, ib, info);
    (void)hipEventRecord(ev.b, stream);
    (void)hipEventSynchronize(ev.b);
    (void)hipEventElapsedTime(&t, ev.a, ev.b);
    out[3] = t / reps;
    // library transforms
    (void)hipStreamSynchronize(stream);
    (void)hipEventRecord(ev.a, stream);
    for (int i = 0; i < reps; ++i)
      if ((rc = gsk_fft_run_(sim, +1)) != GSK_OK) break;
    (void)hipEventRecord(ev.b, stream);
    (void)hipEventSynchronize(ev.b);
    (void)hipEventElapsedTime(&t, ev.a, ev.b);
    out[0] = t / reps;
    (void)hipEventRecord(ev.a, stream);
    for (int i = 0; i < reps; ++i)
      if ((rc = gsk_fft_run_(sim, -1)) != GSK_OK) break;
    (void)hipEventRecord(ev.b, stream);
    (void)hipEventSynchronize(ev.b);
    (void)hipEventElapsedTime(&t, ev.a, ev.b);
    out[2] = t / reps;
    if (rc != GSK_OK) { cleanup(); return rc; }
  } while (false);
  cleanup();
  if (!err.empty()) return gsk_fail_(GSK_ERR, ("fft_compare: " + err).c_str());
  return GSK_OK;
}
