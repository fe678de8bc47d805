// The stress plugin's transform on the GPU (gsk_stress).
#include "gsk_internal.h"

// ======================================================================== stress (src/dynamics_stress.cu)
extern "C" int gsk_stress(int n, double dx, int kconv, const double* C81, const double* eps, int nvar, double xi,
                          int source, const float* eta, float* stress, float* defect, int device) {
  if (!C81 || !eps || !eta || !stress) return fail(GSK_ERR_ARG, "null argument");
  if (nvar < 1 || nvar > 8) return fail(GSK_ERR_ARG, "nvar out of range (1..8)");
  if (source != 1 && source != 2) return fail(GSK_ERR_ARG, "source must be 1 (eta) or 2 (eta^2)");
  gsk_config c;
  CHK(gsk_config_default(&c, GSK_PHYS_MART));
  c.n = n;
  c.nvar = std::max(6, nvar);  // field slots for the 6 stress components
  c.dx = dx;
  c.kconv = kconv;
  c.device = device;
  gsk_sim* s = nullptr;
  CHK(gsk_create(&c, &s));
  const size_t N3 = nvox(s);
  float* bh = nullptr;
  int rc = GSK_OK;
  auto step = [&](int r) { if (rc == GSK_OK) rc = r; };
  // base strains: single unit entries 00, 01, 02, 11, 12, 22 (src/dynamics_stress.cu:70-75)
  double base[54] = {0};
  const int ent[6] = {0, 1, 2, 4, 5, 8};
  for (int b = 0; b < 6; ++b) base[b * 9 + ent[b]] = 1.0;
  const double kfac = kconv == GSK_K_REFERENCE ? 2.0 * 3.14 * dx / n : 2.0 * kPi / (n * dx);
  BArgs a;
  step(fill_bargs(a, C81, base, 6, eps, nvar, kfac));
  if (rc == GSK_OK && hipMalloc((void**)&bh, (size_t)6 * nvar * s->g.kfs * sizeof(float)) != hipSuccess)
    rc = fail(GSK_ERR_HIP, "hipMalloc failed");
  if (rc == GSK_OK && hipMemcpy(s->fld, eta, (size_t)nvar * N3 * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(GSK_ERR_HIP, "copy in failed");
  if (rc == GSK_OK) {
    hipLaunchKernelGGL(k_btensor_half_rect<>, dim3(grid_1d((long)n * n * (n / 2 + 1))), dim3(256), 0, s->stream, bh,
                       s->g, a);
    if (source == 2)
      hipLaunchKernelGGL(k_square<>, dim3(grid_1d((long)nvar * N3)), dim3(256), 0, s->stream, s->fld, (long)nvar * N3);
    if (hipGetLastError() != hipSuccess) rc = fail(GSK_ERR_HIP, "launch failed");
  }
  step(rc == GSK_OK ? dispatch_n<DR2C>(n, s, (const float*)s->fld, s->kbuf, nvar) : rc);
  if (rc == GSK_OK) {
    hipLaunchKernelGGL(k_contract_rect<>, dim3(grid_1d((long)n * n * (n / 2 + 1))), dim3(256), 0, s->stream, s->kbuf,
                       (const float*)bh, s->g, 6, nvar, (float)(-xi / (double)N3));
    if (hipGetLastError() != hipSuccess) rc = fail(GSK_ERR_HIP, "launch failed");
  }
  step(rc == GSK_OK ? dispatch_n<DC2R>(n, s, s->kbuf, s->fld, 6) : rc);
  step(rc == GSK_OK ? sync(s) : rc);
  if (rc == GSK_OK && hipMemcpy(stress, s->fld, 6 * N3 * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(GSK_ERR_HIP, "copy out failed");
  // defect: 1 where any |eta_v| > 0.9 (src/dynamics_stress.cu:178-185)
  if (rc == GSK_OK && defect)
    for (size_t i = 0; i < N3; ++i) {
      float d = 0.f;
      for (int v = 0; v < nvar; ++v)
        if (std::fabs(eta[(size_t)v * N3 + i]) > 0.9f) d = 1.f;
      defect[i] = d;
    }
  if (bh) (void)hipFree(bh);
  gsk_destroy(s);
  return rc;
}

