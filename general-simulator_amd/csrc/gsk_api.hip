// C ABI of the MI355X phase-field step library (include/gsk.h): handles, fields, steps,
// reductions, timing, noise, the B tensor and the library's 3D transforms.
#include "gsk_internal.h"

std::string& err_text() {
  thread_local std::string e;
  return e;
}

// ======================================================================== C ABI
extern "C" {

const char* gsk_last_error(void) { return err_text().c_str(); }

int gsk_philox4x32(const unsigned int ctr[4], unsigned long long seed, unsigned int out[4]) {
  if (!ctr || !out) return fail(GSK_ERR_ARG, "null argument");
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  for (int i = 0; i < 4; ++i) out[i] = c[i];
  return GSK_OK;
}
const char* gsk_version(void) { return "gsk 0.1 (gfx950)"; }

int gsk_config_default(gsk_config* c, int physics) {
  if (!c) return fail(GSK_ERR_ARG, "null config");
  std::memset(c, 0, sizeof(*c));
  c->n = 64;
  c->physics = physics;
  c->path = GSK_PATH_FUSED32;
  c->kconv = GSK_K_EXACT;
  c->dx = 1.0;
  if (physics == GSK_PHYS_MART) {
    c->nvar = 3;
    c->A = 0.2; c->B = 3.0; c->C = 3.0;
    c->kappa = 1.0; c->L = 1.0; c->dt = 0.01; c->w_el = 1.0;
  } else if (physics == GSK_PHYS_PREC) {
    // src/dynamics_precipitate.cu:21-29 defaults
    c->nvar = 4;
    const double A[8] = {0.0, 54.0, -17.0, 7.0, 2.5, 0.2, 0.2, 0.2};
    for (int i = 0; i < 8; ++i) c->a[i] = A[i];
    c->kappa = 300.0; c->kappa_c = 300.0; c->M = 0.4; c->L = 0.4; c->xi = 400.0;
    c->c1 = 0.1; c->c2 = 0.44; c->dt = 0.1;
  } else {
    return fail(GSK_ERR_ARG, "unknown physics");
  }
  // reference-rule weights (src/dynamics_mart.cu:20-25 HEAD side); unused by the spectral rule
  c->rule = GSK_RULE_SPECTRAL;
  c->w_grad = 2.5; c->w_chem = 1.0; c->w_dis = 1.0;
  c->T0 = 450.0; c->temperature = 450.0;
  return GSK_OK;
}

int gsk_config_reference(gsk_config* c, int physics) {
  CHK(gsk_config_default(c, physics));
  c->rule = GSK_RULE_REFERENCE;
  c->kconv = GSK_K_REFERENCE;
  if (physics == GSK_PHYS_MART) {
    // src/dynamics_mart.cu:20-25 (HEAD), LPC[2], LPC[3] at :41
    c->dt = 0.01; c->w_el = 100000.0; c->noise = 0.0001;
    c->B = 32.05; c->C = 37.5;
  } else {
    // src/dynamics_precipitate.cu:21-29; weightnoise 1 -> amplitude 1e-4 (:203, :244)
    c->noise = 0.0001;
    c->S = 0.0;
  }
  return GSK_OK;
}

int gsk_create(const gsk_config* cfg, gsk_sim** out) {
  if (!cfg || !out) return fail(GSK_ERR_ARG, "null argument");
  *out = nullptr;
  if (!supported_n(cfg->n)) return fail(GSK_ERR_UNSUPPORTED, "unsupported grid size " + std::to_string(cfg->n));
  for (int e : {cfg->ny, cfg->nz})
    if (e != 0 && !supported_n(e)) return fail(GSK_ERR_UNSUPPORTED, "unsupported grid size " + std::to_string(e));
  if (cfg->ny < 0 || cfg->nz < 0 || cfg->dy < 0 || cfg->dz < 0) return fail(GSK_ERR_ARG, "negative grid argument");
  const bool box = box_grid(*cfg);
  if (cfg->physics != GSK_PHYS_MART && cfg->physics != GSK_PHYS_PREC) return fail(GSK_ERR_ARG, "unknown physics");
  if (cfg->path != GSK_PATH_FUSED32 && cfg->path != GSK_PATH_PLAIN64) return fail(GSK_ERR_ARG, "unknown path");
  if (cfg->rule != GSK_RULE_SPECTRAL && cfg->rule != GSK_RULE_REFERENCE) return fail(GSK_ERR_ARG, "unknown rule");
  if (cfg->rule == GSK_RULE_REFERENCE && cfg->path != GSK_PATH_FUSED32)
    return fail(GSK_ERR_UNSUPPORTED, "the reference rules run on the FP32 fused path");
  const int maxv = cfg->physics == GSK_PHYS_MART ? 8 : 7;
  if (cfg->nvar < 1 || cfg->nvar > maxv) return fail(GSK_ERR_ARG, "nvar out of range");
  if (!(cfg->dx > 0)) return fail(GSK_ERR_ARG, "dx must be positive");
  CHK(ensure_device(cfg->device));
  gsk_sim* s = new gsk_sim();
  s->cfg = *cfg;
  s->n = cfg->n;
  s->nx = cfg->n;
  s->ny = cfg->ny > 0 ? cfg->ny : cfg->n;
  s->nz = cfg->nz > 0 ? cfg->nz : cfg->n;
  s->V = cfg->nvar;
  s->F = cfg->physics == GSK_PHYS_MART ? s->V : s->V + 1;
  s->dev = cfg->device;
  int rc = dispatch_n<DAttrs>(s->n);
  if (rc != GSK_OK) { delete s; return rc; }
  auto bail = [&](int code) { gsk_destroy(s); return code; };
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(GSK_ERR_HIP, "hipStreamCreate failed"));
  const int n = s->n, M = n / 2;
  const size_t N3 = nvox(s);
  choose_geometry(s, s->F, (s->cfg.physics == GSK_PHYS_MART && s->F > 4) ? s->F : 1);
  choose_plain(s);
  if (box) s->split = false;
  auto up = [&](auto*& dst, const auto& host) -> int {
    using T = std::remove_reference_t<decltype(*dst)>;
    if (hipMalloc((void**)&dst, host.size() * sizeof(T)) != hipSuccess) return fail(GSK_ERR_HIP, "hipMalloc failed");
    if (hipMemcpy(dst, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
      return fail(GSK_ERR_HIP, "hipMemcpy failed");
    return GSK_OK;
  };
  if ((rc = up(s->twN, twiddles<float2>(n, n))) != GSK_OK) return bail(rc);
  if ((rc = up(s->twM, twiddles<float2>(M, M))) != GSK_OK) return bail(rc);
  if ((rc = up(s->rtn, twiddles<float2>(M, n))) != GSK_OK) return bail(rc);
  if ((rc = up(s->twN64, twiddles<double2>(n, n))) != GSK_OK) return bail(rc);
  auto alloc = [&](auto*& p, size_t count) -> int {
    using T = std::remove_reference_t<decltype(*p)>;
    if (hipMalloc((void**)&p, count * sizeof(T)) != hipSuccess)
      return fail(GSK_ERR_HIP, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
    // zeroed on the handle's stream: the stream is non-blocking, so a memset on the null
    // stream would not be ordered before the handle's kernels (it could land after the B
    // tensor build has written the buffer)
    if (hipMemsetAsync(p, 0, count * sizeof(T), s->stream) != hipSuccess)
      return fail(GSK_ERR_HIP, "hipMemsetAsync failed");
    return GSK_OK;
  };
  const size_t T = (size_t)n * s->g.W, ntile = (size_t)s->g.ntile;
  if (box) {
    if ((rc = box_create(s)) != GSK_OK) return bail(rc);
  } else if (cfg->path == GSK_PATH_FUSED32) {
    if ((rc = alloc(s->fld, (size_t)s->F * N3)) != GSK_OK) return bail(rc);
    if ((rc = alloc(s->kbuf, (size_t)s->F * s->g.kfs)) != GSK_OK) return bail(rc);
    if ((rc = alloc(s->state, ntile * s->F * T)) != GSK_OK) return bail(rc);
    const bool vv = cfg->physics == GSK_PHYS_MART || cfg->rule == GSK_RULE_REFERENCE;
    const int npq = vv ? s->V * (s->V + 1) / 2 : 1;
    if ((rc = alloc(s->Bt, ntile * npq * T)) != GSK_OK) return bail(rc);
    if (cfg->rule == GSK_RULE_REFERENCE && (rc = alloc(s->rbuf, (size_t)s->V * N3)) != GSK_OK) return bail(rc);
    if (cfg->rule == GSK_RULE_REFERENCE && cfg->physics == GSK_PHYS_MART &&
        (rc = alloc(s->fstat, (size_t)kRefBlocks * 4)) != GSK_OK)
      return bail(rc);
  } else if (cfg->physics == GSK_PHYS_MART) {
    const size_t VN = (size_t)s->V * N3;
    if ((rc = alloc(s->eta64, VN)) != GSK_OK) return bail(rc);
    if ((rc = alloc(s->S64, VN)) != GSK_OK) return bail(rc);
    if ((rc = alloc(s->R64, VN)) != GSK_OK) return bail(rc);
    if ((rc = alloc(s->E64, VN)) != GSK_OK) return bail(rc);
    if ((rc = alloc(s->F64, VN)) != GSK_OK) return bail(rc);
    if ((rc = alloc(s->B64, (size_t)s->V * s->V * N3)) != GSK_OK) return bail(rc);
  } else {  // precipitate: fields [c, eta_1..eta_V], forces, transforms; the scalar B(k) of c
    const size_t FN = (size_t)s->F * N3;
    if ((rc = alloc(s->eta64, FN)) != GSK_OK) return bail(rc);
    if ((rc = alloc(s->E64, FN)) != GSK_OK) return bail(rc);
    if ((rc = alloc(s->F64, FN)) != GSK_OK) return bail(rc);
    if ((rc = alloc(s->B64, N3)) != GSK_OK) return bail(rc);
  }
  s->rblocks = std::max<int>(1024, (int)ntile);
  if ((rc = alloc(s->part, (size_t)s->rblocks * 16)) != GSK_OK) return bail(rc);
  if ((rc = alloc(s->rout, 16)) != GSK_OK) return bail(rc);
  if (cfg->path == GSK_PATH_FUSED32 && !box && (rc = alloc(s->dstep, 1)) != GSK_OK) return bail(rc);
  if (hipStreamSynchronize(s->stream) != hipSuccess) return bail(fail(GSK_ERR_HIP, "stream synchronisation failed"));
  build_slots(s);
  *out = s;
  return GSK_OK;
}

int gsk_destroy(gsk_sim* s) {
  if (!s) return GSK_OK;
  (void)hipSetDevice(s->dev);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  adj_free(s->adj);
  s->adj = nullptr;
  box_free(s->bx);
  s->bx = nullptr;
  free_dev(s->twN); free_dev(s->twM); free_dev(s->rtn); free_dev(s->twN64);
  if (s->sharedB) s->Bt = nullptr;
  free_dev(s->fld); free_dev(s->kbuf); free_dev(s->state); free_dev(s->Bt); free_dev(s->dstep);
  free_dev(s->eta64); free_dev(s->S64); free_dev(s->R64); free_dev(s->E64); free_dev(s->F64); free_dev(s->B64);
  free_dev(s->part); free_dev(s->rout);
  free_dev(s->rbuf); free_dev(s->defect); free_dev(s->Kd); free_dev(s->fstat);
  for (auto& e : s->ev) (void)hipEventDestroy(e);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return GSK_OK;
}

}  // extern "C"

int fill_bargs(BArgs& a, const double* C81, const double* e1, int V1, const double* e2, int V2, double kfac) {
  if (V1 < 1 || V1 > 8 || V2 < 1 || V2 > 8) return fail(GSK_ERR_ARG, "eigenstrain count out of range (1..8)");
  std::memset(&a, 0, sizeof(a));
  for (int i = 0; i < 81; ++i) a.C[i] = C81[i];
  auto sig = [&](const double* e, int V, double* s) {
    for (int p = 0; p < V; ++p)
      for (int ij = 0; ij < 9; ++ij) {
        double acc = 0;
        for (int kl = 0; kl < 9; ++kl) acc += C81[ij * 9 + kl] * e[p * 9 + kl];
        s[p * 9 + ij] = acc;
      }
  };
  sig(e1, V1, a.s1);
  sig(e2, V2, a.s2);
  for (int p = 0; p < V1; ++p)
    for (int q = 0; q < V2; ++q) {
      double acc = 0;
      for (int ij = 0; ij < 9; ++ij)
        for (int kl = 0; kl < 9; ++kl) acc += C81[ij * 9 + kl] * e1[p * 9 + ij] * e2[q * 9 + kl];
      a.cee[p * V2 + q] = acc;
    }
  a.V1 = V1;
  a.V2 = V2;
  a.kfac = kfac;
  return GSK_OK;
}

extern "C" {

int gsk_build_btensor(gsk_sim* s, const double* C81, const double* eps, int neps) {
  if (!s || !C81 || !eps) return fail(GSK_ERR_ARG, "null argument");
  const int need = (s->cfg.physics == GSK_PHYS_MART || s->cfg.rule == GSK_RULE_REFERENCE) ? s->V : 1;
  if (neps != need) return fail(GSK_ERR_ARG, "eigenstrain count must be " + std::to_string(need));
  CHK(ensure_device(s->dev));
  const double kfac =
      s->cfg.kconv == GSK_K_REFERENCE ? 2.0 * 3.14 * s->cfg.dx / s->n : 2.0 * kPi / (s->n * s->cfg.dx);
  BArgs a;
  CHK(fill_bargs(a, C81, eps, neps, eps, neps, kfac));
  s->cee0 = a.cee[0];
  s->bargs = a;
  s->haveBargs = true;
  for (int i = 0; i < neps * 9; ++i) s->eps[i] = eps[i];
  s->haveEps = true;
  if (s->bx) {
    CHK(box_build_btensor(s, a));
  } else if (s->cfg.path == GSK_PATH_FUSED32 && s->split && s->cfg.physics == GSK_PHYS_PREC &&
             s->cfg.rule == GSK_RULE_SPECTRAL) {
    // the concentration's scalar B(k), read by the precipitate's x update (its tile order)
    const long total = (long)s->sg.ntile * s->n * split_w<kSplitWc>();
    hipLaunchKernelGGL(k_btensor_split<kSplitWc>, dim3(grid_1d(total)), dim3(256), 0, s->stream, s->Bt, s->sg, a);
  } else if (s->cfg.path == GSK_PATH_FUSED32 && s->split) {  // read by the contraction
    const long total = (long)s->sg.ntile_c * s->n * split_w<kSplitWcC>();
    hipLaunchKernelGGL(k_btensor_split<kSplitWcC>, dim3(grid_1d(total)), dim3(256), 0, s->stream, s->Bt, s->sg, a);
  } else if (s->cfg.path == GSK_PATH_FUSED32) {
    const long total = (long)(long)s->g.ntile * s->n * s->g.W;
    hipLaunchKernelGGL(k_btensor_tiles<>, dim3(grid_1d(total)), dim3(256), 0, s->stream, s->Bt, s->g, a);
  } else {
    hipLaunchKernelGGL(k_btensor_full<>, dim3(grid_1d((long)nvox(s))), dim3(256), 0, s->stream, s->B64, s->n, a);
  }
  HIPCHK(hipGetLastError());
  CHK(sync(s));
  s->haveB = true;
  s->dirty = true;
  ++s->gen;
  return GSK_OK;
}

static int field_info(gsk_sim* s, const char* name, int* offset_fields, int* count) {
  if (!name) return fail(GSK_ERR_ARG, "null field name");
  if (std::strcmp(name, "eta") == 0) {
    *offset_fields = s->cfg.physics == GSK_PHYS_MART ? 0 : 1;
    *count = s->V;
    return GSK_OK;
  }
  if (std::strcmp(name, "concentration") == 0 && s->cfg.physics == GSK_PHYS_PREC) {
    *offset_fields = 0;
    *count = 1;
    return GSK_OK;
  }
  return fail(GSK_ERR_ARG, std::string("unknown field \"") + name + "\"");
}

static int set_ref_field(gsk_sim* s, const char* name, const void* host, int dtype) {
  const size_t N3 = nvox(s);
  auto val = [&](size_t i) { return dtype == GSK_F32 ? (double)((const float*)host)[i] : ((const double*)host)[i]; };
  if (s->cfg.rule != GSK_RULE_REFERENCE || s->cfg.physics != GSK_PHYS_MART)
    return fail(GSK_ERR_UNSUPPORTED, std::string("field \"") + name + "\" belongs to the martensite reference rule");
  std::vector<float> h;
  float** dst;
  if (std::strcmp(name, "defect") == 0) {
    h.resize(N3);
    for (size_t i = 0; i < N3; ++i) h[i] = (float)val(i);
    dst = &s->defect;
  } else {
    // sigma from the 6 stored components (src/dynamics_mart.cu:125-139), then
    // K_p = 25 sum_ij sigma_ij eps^p_ij (src/dynamics_mart.cu:305-318)
    if (!s->haveEps) return fail(GSK_ERR_STATE, "build the B tensor before setting the dislocation stress");
    static const int comp[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
    h.resize((size_t)s->V * N3);
    for (int p = 0; p < s->V; ++p)
      for (size_t i = 0; i < N3; ++i) {
        double k = 0.0;
        for (int ij = 0; ij < 9; ++ij) k += 25.0 * val(comp[ij] * N3 + i) * s->eps[p * 9 + ij];
        h[(size_t)p * N3 + i] = (float)k;
      }
    dst = &s->Kd;
  }
  CHK(sync(s));
  if (!*dst) HIPCHK(hipMalloc((void**)dst, h.size() * sizeof(float)));
  HIPCHK(hipMemcpy(*dst, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
  ++s->gen;
  build_slots(s);  // the update now reads this field too
  return GSK_OK;
}

int gsk_set_field(gsk_sim* s, const char* name, const void* host, int dtype) {
  if (!s || !host) return fail(GSK_ERR_ARG, "null argument");
  if (dtype != GSK_F32 && dtype != GSK_F64) return fail(GSK_ERR_ARG, "bad dtype");
  if (name && (std::strcmp(name, "defect") == 0 || std::strcmp(name, "dislocationstress") == 0)) {
    CHK(ensure_device(s->dev));
    return set_ref_field(s, name, host, dtype);
  }
  int off, cnt;
  CHK(field_info(s, name, &off, &cnt));
  CHK(ensure_device(s->dev));
  const size_t N3 = nvox(s), count = (size_t)cnt * N3;
  if (s->bx) {
    std::vector<double> h(count);
    for (size_t i = 0; i < count; ++i)
      h[i] = dtype == GSK_F32 ? (double)((const float*)host)[i] : ((const double*)host)[i];
    return box_set_field(s, off, count, h.data());
  }
  if (s->cfg.path == GSK_PATH_PLAIN64) {
    std::vector<double> h(count);
    for (size_t i = 0; i < count; ++i)
      h[i] = dtype == GSK_F32 ? (double)((const float*)host)[i] : ((const double*)host)[i];
    CHK(sync(s));  // earlier work on the handle's stream may still read the field
    HIPCHK(hipMemcpy(s->eta64 + (size_t)off * N3, h.data(), count * sizeof(double), hipMemcpyHostToDevice));
    return GSK_OK;
  }
  std::vector<float> h(count);
  const bool conc = std::strcmp(name, "concentration") == 0;
  if (conc) {
    // keep the mean in FP64 (compensated sum) and store the fluctuation in FP32;
    // the state's k = 0 mode of the fluctuation is held at exactly zero
    double sum = 0.0, comp = 0.0;
    for (size_t i = 0; i < count; ++i) {
      const double v = dtype == GSK_F32 ? (double)((const float*)host)[i] : ((const double*)host)[i];
      const double y = v - comp, t = sum + y;
      comp = (t - sum) - y;
      sum = t;
    }
    // The mean is rounded to binary32 so that v - cbar is exact in FP32 for inputs near
    // the mean (no correlated rounding of the fluctuations); the FP64 sum of the
    // fluctuations becomes the k = 0 mode of the state.
    const double cbar = (double)(float)(sum / (double)count);
    double k0 = 0.0, ck = 0.0;
    for (size_t i = 0; i < count; ++i) {
      const double v = dtype == GSK_F32 ? (double)((const float*)host)[i] : ((const double*)host)[i];
      h[i] = (float)(v - cbar);
      const double y = (double)h[i] - ck, t = k0 + y;
      ck = (t - k0) - y;
      k0 = t;
    }
    s->cbar = cbar;
    s->c_k0 = k0;
  } else {
    for (size_t i = 0; i < count; ++i) h[i] = dtype == GSK_F32 ? ((const float*)host)[i] : (float)((const double*)host)[i];
  }
  CHK(sync(s));  // earlier work on the handle's stream may still read the field
  HIPCHK(hipMemcpy(s->fld + (size_t)off * N3, h.data(), count * sizeof(float), hipMemcpyHostToDevice));
  s->dirty = true;
  ++s->gen;
  return GSK_OK;
}

int gsk_get_field(gsk_sim* s, const char* name, void* host, int dtype) {
  if (!s || !host) return fail(GSK_ERR_ARG, "null argument");
  if (dtype != GSK_F32 && dtype != GSK_F64) return fail(GSK_ERR_ARG, "bad dtype");
  int off, cnt;
  CHK(field_info(s, name, &off, &cnt));
  CHK(ensure_device(s->dev));
  CHK(sync(s));
  const size_t N3 = nvox(s), count = (size_t)cnt * N3;
  if (s->bx) {
    std::vector<double> h(count);
    CHK(box_get_field(s, off, count, h.data()));
    for (size_t i = 0; i < count; ++i) {
      if (dtype == GSK_F32) ((float*)host)[i] = (float)h[i];
      else ((double*)host)[i] = h[i];
    }
    return GSK_OK;
  }
  if (s->cfg.path == GSK_PATH_PLAIN64) {
    std::vector<double> h(count);
    HIPCHK(hipMemcpy(h.data(), s->eta64 + (size_t)off * N3, count * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < count; ++i) {
      if (dtype == GSK_F32) ((float*)host)[i] = (float)h[i];
      else ((double*)host)[i] = h[i];
    }
    return GSK_OK;
  }
  std::vector<float> h(count);
  HIPCHK(hipMemcpy(h.data(), s->fld + (size_t)off * N3, count * sizeof(float), hipMemcpyDeviceToHost));
  const double add = std::strcmp(name, "concentration") == 0 ? s->cbar : 0.0;
  for (size_t i = 0; i < count; ++i) {
    if (dtype == GSK_F32) ((float*)host)[i] = (float)(add + (double)h[i]);
    else ((double*)host)[i] = add + (double)h[i];
  }
  return GSK_OK;
}

int gsk_init_uniform(gsk_sim* s, const char* name, double lo, double hi, unsigned long long seed) {
  if (!s) return fail(GSK_ERR_ARG, "null argument");
  int off, cnt;
  CHK(field_info(s, name, &off, &cnt));
  CHK(ensure_device(s->dev));
  const size_t N3 = nvox(s);
  // generated on the device in FP32, then passed through set_field so the
  // concentration mean handling and the FP64 path see the same values
  float* tmp = nullptr;
  HIPCHK(hipMalloc((void**)&tmp, (size_t)cnt * N3 * sizeof(float)));
  for (int f = 0; f < cnt; ++f)
    hipLaunchKernelGGL(k_uniform<>, dim3(grid_1d((long)N3 / 4)), dim3(256), 0, s->stream, tmp + f * N3, (long)N3, seed,
                       (uint32_t)f, lo, hi);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  std::vector<float> h((size_t)cnt * N3);
  if (e == hipSuccess) e = hipMemcpy(h.data(), tmp, h.size() * sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(tmp);
  if (e != hipSuccess) return fail(GSK_ERR_HIP, std::string("init_uniform: ") + hipGetErrorString(e));
  return gsk_set_field(s, name, h.data(), GSK_F32);
}

int gsk_step(gsk_sim* s, int nsteps) {
  if (!s) return fail(GSK_ERR_ARG, "null handle");
  if (nsteps < 0) return fail(GSK_ERR_ARG, "negative step count");
  CHK(ensure_device(s->dev));
  CHK(prepare(s));
  for (int i = 0; i < nsteps; ++i) {
    CHK(step_once(s));
    ++s->step;
  }
  return sync(s);
}

int gsk_set_step(gsk_sim* s, long long step) {
  if (!s || step < 0) return fail(GSK_ERR_ARG, "bad argument");
  s->step = step;
  s->sq_ready = false;  // carried-over noise was drawn for the old step index
  if (s->dstep) {
    CHK(ensure_device(s->dev));
    CHK(sync(s));
    const uint32_t v = (uint32_t)step;
    HIPCHK(hipMemcpy(s->dstep, &v, sizeof(v), hipMemcpyHostToDevice));
  }
  return GSK_OK;
}
long long gsk_get_step(const gsk_sim* s) { return s ? s->step : -1; }

int gsk_set_param(gsk_sim* s, const char* key, double v) {
  if (!s || !key) return fail(GSK_ERR_ARG, "null argument");
  gsk_config& c = s->cfg;
  std::string k(key);
  if (k == "A") c.A = v; else if (k == "B") c.B = v; else if (k == "C") c.C = v;
  else if (k == "L") c.L = v; else if (k == "kappa") c.kappa = v; else if (k == "dt") c.dt = v;
  else if (k == "noise") c.noise = v; else if (k == "w_el") c.w_el = v; else if (k == "M") c.M = v;
  else if (k == "kappa_c") c.kappa_c = v; else if (k == "xi") c.xi = v; else if (k == "c1") c.c1 = v;
  else if (k == "c2") c.c2 = v; else if (k == "S") c.S = v;
  else if (k == "w_grad") c.w_grad = v; else if (k == "w_chem") c.w_chem = v; else if (k == "w_dis") c.w_dis = v;
  else if (k == "temperature") c.temperature = v; else if (k == "T0") c.T0 = v;
  else if (k == "chem_form") c.chem_form = (int)v;
  else if (k.size() == 2 && k[0] == 'a' && k[1] >= '1' && k[1] <= '7') c.a[k[1] - '0'] = v;
  else return fail(GSK_ERR_ARG, "unknown parameter " + k);
  ++s->gen;
  // the precipitate's carried-over local forces depend on the parameters
  if (s->cfg.physics == GSK_PHYS_PREC) s->sq_ready = false;
  return GSK_OK;
}

int gsk_reduce(gsk_sim* s, int what, double* out) {
  if (!s || !out) return fail(GSK_ERR_ARG, "null argument");
  CHK(ensure_device(s->dev));
  CHK(prepare(s));
  const long N3 = (long)nvox(s);
  const int nb = 512;
  const gsk_config& c = s->cfg;
  double h[16];
  // the real fields in FP64 (plain path, FP64 box) or FP32 (fused path, FP32 box)
  const double* f64 = s->bx ? (s->bx->f64 ? (const double*)s->bx->fld : nullptr) : s->eta64;
  const float* f32 = s->bx ? (s->bx->f64 ? nullptr : (const float*)s->bx->fld) : s->fld;
  const bool plain = f64 != nullptr;
  auto real_pass = [&]() -> int {
    if (c.physics == GSK_PHYS_MART) {
      if (plain) LAUNCH(r_mart_real<double>, nb, 256, 0, s, f64, N3, s->V, c.A, c.B, c.C, s->part);
      else LAUNCH(r_mart_real<float>, nb, 256, 0, s, f32, N3, s->V, c.A, c.B, c.C, s->part);
    } else {
      PrecE pe{};
      for (int i = 0; i < 8; ++i) pe.A[i] = c.a[i];
      pe.c1 = c.c1; pe.c2 = c.c2; pe.cbar = plain ? 0.0 : s->cbar; pe.V = s->V;
      if (plain) LAUNCH(r_prec_real<double>, nb, 256, 0, s, f64, N3, pe, s->part);
      else LAUNCH(r_prec_real<float>, nb, 256, 0, s, f32, N3, pe, s->part);
    }
    LAUNCH(r_finish<>, 1, 1024, 0, s, s->part, nb, 2 + s->V, s->rout);
    HIPCHK(hipMemcpyAsync(h, s->rout, (2 + s->V) * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    return sync(s);
  };
  if (what == GSK_RED_MEAN_ETA || what == GSK_RED_SUM_ETA2) {
    CHK(real_pass());
    if (what == GSK_RED_SUM_ETA2) {
      if (c.physics != GSK_PHYS_MART) return fail(GSK_ERR_UNSUPPORTED, "sum eta^2 is a martensite reduction");
      out[0] = h[1];
    } else {
      for (int v = 0; v < s->V; ++v) out[v] = h[2 + v] / (double)N3;
    }
    return GSK_OK;
  }
  if (what == GSK_RED_MART_FORCES) {
    if (c.physics != GSK_PHYS_MART || c.rule != GSK_RULE_REFERENCE || !s->fstat)
      return fail(GSK_ERR_UNSUPPORTED, "the force means belong to the martensite reference rule");
    LAUNCH(r_finish<>, 1, 1024, 0, s, s->fstat, s->fblocks, 4, s->rout);
    HIPCHK(hipMemcpyAsync(h, s->rout, 4 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    CHK(sync(s));
    for (int j = 0; j < 4; ++j) out[j] = h[j] / ((double)s->V * (double)N3);
    return GSK_OK;
  }
  if (what == GSK_RED_MASS) {
    if (c.physics != GSK_PHYS_PREC) return fail(GSK_ERR_UNSUPPORTED, "mass is a precipitate reduction");
    CHK(prepare(s));
    if (s->bx || s->cfg.path == GSK_PATH_PLAIN64) {  // the FP64 sum of the stored field (+ N cbar), twice
      CHK(real_pass());
      out[0] = out[1] = (s->bx ? s->cbar * (double)N3 : 0.0) + h[1];
      return GSK_OK;
    }
    // out[0]: N * mean + Re of the k = 0 mode of the fluctuation state (the conserved quantity)
    // out[1]: FP64 sum of the FP32 real field (includes one IFFT's rounding, not accumulated)
    float2 k0 = make_float2(0.f, 0.f);
    HIPCHK(hipMemcpyAsync(&k0, s->state + k0_index(s), sizeof(float2), hipMemcpyDeviceToHost, s->stream));
    CHK(sync(s));
    CHK(real_pass());
    out[0] = s->cbar * (double)N3 + (double)k0.x;
    out[1] = s->cbar * (double)N3 + h[1];
    return GSK_OK;
  }
  if (what == GSK_RED_ENERGY) {
    if (s->cfg.rule == GSK_RULE_REFERENCE)
      return fail(GSK_ERR_UNSUPPORTED, "the reference rules define no free energy (the plugins report forces)");
    CHK(real_pass());
    const double chem = h[0];
    double gr = 0.0, el = 0.0;
    if (s->bx) {
      CHK(box_energy(s, &gr, &el));
    } else if (s->cfg.path == GSK_PATH_PLAIN64 && c.physics == GSK_PHYS_PREC) {
      double ks[3];
      CHK(dispatch_n<DEpP>(s->n, s, ks));
      gr = 0.5 * (c.kappa_c * ks[0] + c.kappa * ks[1]) / (double)N3;
      el = 0.5 * c.xi * ks[2] / (double)N3;
    } else if (s->cfg.path == GSK_PATH_PLAIN64) {
      double g2, e2;
      CHK(dispatch_n<DElP>(s->n, s, &g2, &e2));
      gr = 0.5 * c.kappa * g2 / (double)N3;
      el = 0.5 * c.w_el * e2 / (double)N3;
    } else {
      const int ntile = s->g.ntile;
      const double kfac = c.kconv == GSK_K_REFERENCE ? 2.0 * 3.14 * c.dx / s->n : 2.0 * kPi / (s->n * c.dx);
      if (s->split) LAUNCH(r_state_grad_split<>, nb, 256, 0, s, s->state, s->sg, s->F, (float)kfac, s->part);
      else LAUNCH(r_state_grad<>, nb, 256, 0, s, s->state, s->g, s->F, (float)kfac, s->part);
      LAUNCH(r_finish<>, 1, 1024, 0, s, s->part, nb, s->F, s->rout);
      double gh[16];
      HIPCHK(hipMemcpyAsync(gh, s->rout, s->F * sizeof(double), hipMemcpyDeviceToHost, s->stream));
      CHK(sync(s));
      if (c.physics == GSK_PHYS_MART) {
        double acc = 0;
        for (int f = 0; f < s->F; ++f) acc += gh[f];
        gr = 0.5 * c.kappa * acc / (double)N3;
        double e2;
        CHK(dispatch_n<DElF>(s->n, s, &e2));
        el = 0.5 * c.w_el * e2 / (double)N3;
      } else {
        double acc = 0;
        for (int f = 1; f < s->F; ++f) acc += gh[f];
        gr = 0.5 * c.kappa_c * gh[0] / (double)N3 + 0.5 * c.kappa * acc / (double)N3;
        if (s->split) LAUNCH(r_state_el_prec_split<>, nb, 256, 0, s, s->state, s->Bt, s->sg, s->F, s->part);
        else LAUNCH(r_state_el_prec<>, nb, 256, 0, s, s->state, s->Bt, s->g, s->F, s->part);
        LAUNCH(r_finish<>, 1, 1024, 0, s, s->part, nb, 1, s->rout);
        double eh;
        HIPCHK(hipMemcpyAsync(&eh, s->rout, sizeof(double), hipMemcpyDeviceToHost, s->stream));
        CHK(sync(s));
        // the k = 0 mode of c is N * cbar + the state's k = 0 mode of the fluctuation
        float2 k0 = make_float2(0.f, 0.f);
        HIPCHK(hipMemcpyAsync(&k0, s->state + k0_index(s), sizeof(float2), hipMemcpyDeviceToHost, s->stream));
    CHK(sync(s));
        const double f0 = (double)k0.x, c0 = (double)N3 * s->cbar + f0;
        el = 0.5 * c.xi * (eh - s->cee0 * f0 * f0 + s->cee0 * c0 * c0) / (double)N3;
        (void)ntile;
      }
    }
    out[0] = chem; out[1] = gr; out[2] = el; out[3] = chem + gr + el;
    return GSK_OK;
  }
  return fail(GSK_ERR_ARG, "unknown reduction");
}

int gsk_time_steps(gsk_sim* s, int nsteps, double* ms) {
  if (!s || !ms || nsteps <= 0) return fail(GSK_ERR_ARG, "bad argument");
  CHK(ensure_device(s->dev));
  CHK(prepare(s));
  hipEvent_t a = nullptr, b = nullptr;
  struct Events {  // destroyed on every return path
    hipEvent_t* a;
    hipEvent_t* b;
    ~Events() {
      if (*a) (void)hipEventDestroy(*a);
      if (*b) (void)hipEventDestroy(*b);
    }
  } guard{&a, &b};
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  HIPCHK(hipEventRecord(a, s->stream));
  for (int i = 0; i < nsteps; ++i) {
    CHK(step_once(s));
    ++s->step;
  }
  HIPCHK(hipEventRecord(b, s->stream));
  HIPCHK(hipEventSynchronize(b));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, a, b));
  *ms = (double)t;
  return GSK_OK;
}

int gsk_kernel_times(gsk_sim* s, int nsteps, double* ms, int maxk, int* nk, char* names, int names_len) {
  if (!s || !ms || !nk || nsteps <= 0) return fail(GSK_ERR_ARG, "bad argument");
  CHK(ensure_device(s->dev));
  CHK(prepare(s));
  const int ns = (int)s->slots.size();
  const int need = ns * nsteps + 1;
  while ((int)s->ev.size() < need) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    s->ev.push_back(e);
  }
  if (s->bx) return fail(GSK_ERR_UNSUPPORTED, "per-kernel timing covers the cube schedules");
  s->prof = true;
  s->evi = 0;
  for (int i = 0; i < nsteps; ++i) {
    int rc = dispatch_n<DStep>(s->n, s);
    if (rc != GSK_OK) { s->prof = false; return rc; }
    ++s->step;
  }
  CHK(mark(s));
  s->prof = false;
  HIPCHK(hipEventSynchronize(s->ev[s->evi - 1]));
  if (s->evi != need) return fail(GSK_ERR, "kernel slot count mismatch");
  const int out = std::min(ns, maxk);
  for (int k = 0; k < out; ++k) ms[k] = 0.0;
  for (int i = 0; i < nsteps; ++i)
    for (int k = 0; k < ns; ++k) {
      float t = 0.f;
      HIPCHK(hipEventElapsedTime(&t, s->ev[i * ns + k], s->ev[i * ns + k + 1]));
      if (k < out) ms[k] += t;
    }
  *nk = ns;
  if (names && names_len > 0) {
    std::string all;
    for (int k = 0; k < ns; ++k) { if (k) all += ","; all += s->slots[k].name; }
    std::snprintf(names, names_len, "%s", all.c_str());
  }
  return GSK_OK;
}

int gsk_kernel_bytes(gsk_sim* s, double* bytes, int maxk, int* nk) {
  if (!s || !bytes || !nk) return fail(GSK_ERR_ARG, "null argument");
  const int ns = (int)s->slots.size();
  for (int k = 0; k < std::min(ns, maxk); ++k) bytes[k] = s->slots[k].bytes;
  *nk = ns;
  return GSK_OK;
}

int gsk_noise_field(unsigned long long seed, unsigned int step, unsigned int field, long long nv, float* host_out,
                    int device) {
  if (!host_out || nv <= 0 || nv % 4) return fail(GSK_ERR_ARG, "nvox must be a positive multiple of 4");
  CHK(ensure_device(device));
  float* d = nullptr;
  HIPCHK(hipMalloc((void**)&d, nv * sizeof(float)));
  hipLaunchKernelGGL(k_noise<>, dim3(grid_1d(nv / 4)), dim3(256), 0, 0, d, (long)nv, seed, step, field);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpy(host_out, d, nv * sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(GSK_ERR_HIP, std::string("noise_field: ") + hipGetErrorString(e));
  return GSK_OK;
}

int gsk_btensor_full(int n, double dx, int kconv, const double* C81, const double* eps1, int V1, const double* eps2,
                     int V2, double* host_out, int device) {
  if (!C81 || !eps1 || !eps2 || !host_out || n < 2) return fail(GSK_ERR_ARG, "bad argument");
  CHK(ensure_device(device));
  const double kfac = kconv == GSK_K_REFERENCE ? 2.0 * 3.14 * dx / n : 2.0 * kPi / (n * dx);
  BArgs a;
  CHK(fill_bargs(a, C81, eps1, V1, eps2, V2, kfac));
  const size_t N3 = (size_t)n * n * n, cnt = N3 * V1 * V2;
  double* d = nullptr;
  HIPCHK(hipMalloc((void**)&d, cnt * sizeof(double)));
  hipLaunchKernelGGL(k_btensor_full<>, dim3(grid_1d((long)N3)), dim3(256), 0, 0, d, n, a);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpy(host_out, d, cnt * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(GSK_ERR_HIP, std::string("btensor_full: ") + hipGetErrorString(e));
  return GSK_OK;
}

static int fft_sim(int n, int count, int device, gsk_sim** out) {
  if (count < 1 || count > 8) return fail(GSK_ERR_ARG, "count out of range (1..8)");
  gsk_config c;
  CHK(gsk_config_default(&c, GSK_PHYS_MART));
  c.n = n;
  c.nvar = count;
  c.device = device;
  return gsk_create(&c, out);
}

int gsk_fail_(int code, const char* msg) { return fail(code, msg ? msg : "error"); }

int gsk_fft_prepare_(int n, int count, int device, void** sim, float** real, float2** kbuf, hipStream_t* stream) {
  gsk_sim* s = nullptr;
  CHK(fft_sim(n, count, device, &s));
  *sim = s;
  *real = s->fld;
  *kbuf = s->kbuf;
  *stream = s->stream;
  return GSK_OK;
}

// dir > 0: R2C of the real fields into the half spectrum; dir < 0: C2R back (unnormalised)
int gsk_fft_run_(void* sim, int dir) {
  gsk_sim* s = static_cast<gsk_sim*>(sim);
  const int n = s->n, F = s->V;
  if (dir > 0) {
    CHK(dispatch_n<DZfwd>(n, s, (const float*)s->fld, F));
    CHK(dispatch_n<DY>(n, s, true, F));
    return dispatch_n<DX>(n, s, true, F, 1.0f);
  }
  CHK(dispatch_n<DX>(n, s, false, F, 1.0f));
  CHK(dispatch_n<DY>(n, s, false, F));
  return dispatch_n<DZinv>(n, s, s->fld, F);
}

int gsk_fft_release_(void* sim) { return gsk_destroy(static_cast<gsk_sim*>(sim)); }

// Copy the half spectra of the handle's fields to host in the compact [f][kx][ky][kz] layout.
int gsk_fft_spectrum_(void* sim, float* out) {
  gsk_sim* s = static_cast<gsk_sim*>(sim);
  const int n = s->n, M = n / 2, count = s->V;
  CHK(sync(s));
  std::vector<float2> h((size_t)count * s->g.kfs);
  HIPCHK(hipMemcpy(h.data(), s->kbuf, h.size() * sizeof(float2), hipMemcpyDeviceToHost));
  for (size_t f = 0; f < (size_t)count; ++f)
    for (size_t r = 0; r < (size_t)n * n; ++r)
      for (size_t k = 0; k <= (size_t)M; ++k) {
        const float2 v = h[kidx(s->g, (int)f, (long long)r, (int)k)];
        const size_t o = ((f * n * n + r) * (M + 1) + k) * 2;
        out[o] = v.x;
        out[o + 1] = v.y;
      }
  return GSK_OK;
}

int gsk_fft_r2c(int n, int count, const float* in, float* out, int device) {
  if (!in || !out) return fail(GSK_ERR_ARG, "null argument");
  gsk_sim* s = nullptr;
  CHK(fft_sim(n, count, device, &s));
  const size_t N3 = nvox(s), M = n / 2;
  int rc = GSK_OK;
  if (hipMemcpy(s->fld, in, count * N3 * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(GSK_ERR_HIP, "copy in failed");
  if (rc == GSK_OK) rc = dispatch_n<DZfwd>(n, s, (const float*)s->fld, count);
  if (rc == GSK_OK) rc = dispatch_n<DY>(n, s, true, count);
  if (rc == GSK_OK) rc = dispatch_n<DX>(n, s, true, count, 1.0f);
  if (rc == GSK_OK) rc = sync(s);
  if (rc == GSK_OK) rc = gsk_fft_spectrum_(s, out);
  (void)M;
  gsk_destroy(s);
  return rc;
}

int gsk_fft_c2r(int n, int count, const float* in, float* out, int device) {
  if (!in || !out) return fail(GSK_ERR_ARG, "null argument");
  gsk_sim* s = nullptr;
  CHK(fft_sim(n, count, device, &s));
  const size_t N3 = nvox(s), M = n / 2;
  std::vector<float2> h((size_t)count * s->g.kfs, make_float2(0.f, 0.f));
  for (size_t f = 0; f < (size_t)count; ++f)
    for (size_t r = 0; r < (size_t)n * n; ++r)
      for (size_t k = 0; k <= M; ++k) {
        size_t o = ((f * n * n + r) * (M + 1) + k) * 2;
        h[kidx(s->g, f, r, k)] = make_float2(in[o], in[o + 1]);
      }
  int rc = GSK_OK;
  if (hipMemcpy(s->kbuf, h.data(), h.size() * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(GSK_ERR_HIP, "copy in failed");
  if (rc == GSK_OK) rc = dispatch_n<DX>(n, s, false, count, 1.0f);
  if (rc == GSK_OK) rc = dispatch_n<DY>(n, s, false, count);
  if (rc == GSK_OK) rc = dispatch_n<DZinv>(n, s, s->fld, count);
  if (rc == GSK_OK) rc = sync(s);
  if (rc == GSK_OK && hipMemcpy(out, s->fld, count * N3 * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(GSK_ERR_HIP, "copy out failed");
  gsk_destroy(s);
  return rc;
}

}  // extern "C"