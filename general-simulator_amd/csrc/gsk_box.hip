// Box grids: nx x ny x nz voxels with spacings dx, dy, dz (the reference's `gridsize nx ny nz dx dy
// dz`, src/dynamics_mart.cu:18), for any edges out of the supported transform lengths.
//
// The reference builds its wave numbers per axis from that axis's n and spacing
// (src/gtensorb.cu:14-24) and divides each difference quotient of its gradient term by that axis's
// spacing (src/dynamics_mart.cu:181-188). A box runs the plain-path kernels (one kernel per
// operation, full complex spectra) with 3D transforms as three axis passes of the LDS Stockham
// engine, each pass compiled for its own axis length: in FP32 for GSK_PATH_FUSED32 (both rules)
// and in FP64 for GSK_PATH_PLAIN64 (the spectral rules). The fields are the state (no k-space state
// is carried between steps), stored [f][x][y][z] like the reference's Data<Real>.
#include "gsk_internal.h"

namespace {

template <typename T> using C2 = typename Cx<T>::T2;

// kernels of the transform passes of one axis length N, both precisions and directions
template <int N> struct BoxN {
  static int attrs() {
    CHK(lds_limit((const void*)p_zrows<float2, N, true>));
    CHK(lds_limit((const void*)p_zrows<float2, N, false>));
    CHK(lds_limit((const void*)p_strided<float2, N, true>));
    CHK(lds_limit((const void*)p_strided<float2, N, false>));
    CHK(lds_limit((const void*)p_zrows<double2, N, true>));
    CHK(lds_limit((const void*)p_zrows<double2, N, false>));
    CHK(lds_limit((const void*)p_strided<double2, N, true>));
    CHK(lds_limit((const void*)p_strided<double2, N, false>));
    return GSK_OK;
  }
  // rows along z (N = nz): nrows rows in tiles of L
  template <typename T2> static int zpass(gsk_sim* s, T2* b, bool fwd) {
    const BoxWS& w = *s->bx;
    const int gz = (int)((long)s->nx * s->ny / w.L);
    if (fwd) LAUNCH((p_zrows<T2, N, true>), gz, w.tz, w.ldz, s, b, w.L, (const T2*)w.tw[2]);
    else LAUNCH((p_zrows<T2, N, false>), gz, w.tz, w.ldz, s, b, w.L, (const T2*)w.tw[2]);
    return GSK_OK;
  }
  // strided lines along y (axis 1: outer x, stride ny nz; positions stride nz) or x (axis 0: outer y,
  // stride nz; positions stride ny nz), the nz columns in chunks of W
  template <typename T2> static int spass(gsk_sim* s, T2* b, int axis, bool fwd) {
    const BoxWS& w = *s->bx;
    const long nyz = (long)s->ny * s->nz;
    const int W = w.W[axis], nch = s->nz / W, outer = axis == 1 ? s->nx : s->ny;
    const long os = axis == 1 ? nyz : (long)s->nz, ps = axis == 1 ? (long)s->nz : nyz;
    const int th = axis == 1 ? w.ty : w.tx;
    const size_t lds = axis == 1 ? w.ldy : w.ldx;
    const T2* tw = (const T2*)w.tw[axis];
    if (fwd) LAUNCH((p_strided<T2, N, true>), outer * nch, th, lds, s, b, W, nch, os, ps, tw);
    else LAUNCH((p_strided<T2, N, false>), outer * nch, th, lds, s, b, W, nch, os, ps, tw);
    return GSK_OK;
  }
};

template <typename T2> int axis_pass(gsk_sim* s, T2* b, int axis, bool fwd) {
  const int n = axis == 0 ? s->nx : (axis == 1 ? s->ny : s->nz);
#define GSK_BOX_AXIS(N) \
  case N: return axis == 2 ? BoxN<N>::template zpass<T2>(s, b, fwd) : BoxN<N>::template spass<T2>(s, b, axis, fwd);
  switch (n) {
    GSK_BOX_AXIS(16) GSK_BOX_AXIS(32) GSK_BOX_AXIS(64) GSK_BOX_AXIS(96)
    GSK_BOX_AXIS(128) GSK_BOX_AXIS(192) GSK_BOX_AXIS(256) GSK_BOX_AXIS(512)
  }
#undef GSK_BOX_AXIS
  return fail(GSK_ERR_UNSUPPORTED, "unsupported axis length " + std::to_string(n));
}

// unnormalised 3D C2C of nf fields of N3 values each
template <typename T2> int c2c(gsk_sim* s, T2* buf, bool fwd, int nf) {
  const long N3 = (long)nvox(s);
  for (int f = 0; f < nf; ++f)
    for (int axis = 2; axis >= 0; --axis) CHK(axis_pass<T2>(s, buf + f * N3, axis, fwd));
  return GSK_OK;
}

double kf_of(const gsk_config& c, int n, double d) {
  return c.kconv == GSK_K_REFERENCE ? 2.0 * 3.14 * d / n : 2.0 * kPi / (n * d);
}

template <typename T> PlainKT<T> mart_k(const gsk_sim* s) {
  const gsk_config& c = s->cfg;
  const BoxWS& w = *s->bx;
  PlainKT<T> k{};
  k.A = (T)c.A; k.B = (T)c.B; k.C = (T)c.C;
  k.two_wel = (T)(2.0 * c.w_el);
  k.noise = (T)c.noise;
  k.dtL = (T)(c.dt * c.L);
  k.dtLk = (T)(c.dt * c.L * c.kappa);
  k.kfx = (T)kf_of(c, s->nx, w.dx); k.kfy = (T)kf_of(c, s->ny, w.dy); k.kfz = (T)kf_of(c, s->nz, w.dz);
  k.d = BoxDim{s->nx, s->ny, s->nz};
  k.inv_n3 = (T)(1.0 / (double)nvox(s));
  k.V = s->V;
  k.step = (uint32_t)s->step;
  k.seed = c.seed;
  return k;
}

template <typename T> PlainPKT<T> prec_k(const gsk_sim* s) {
  const gsk_config& c = s->cfg;
  const BoxWS& w = *s->bx;
  const bool ref = c.rule == GSK_RULE_REFERENCE;
  PlainPKT<T> k{};
  for (int i = 0; i < 8; ++i) k.A[i] = (T)c.a[i];
  k.c1 = (T)c.c1; k.c2 = (T)c.c2;
  k.dtM = (T)(c.dt * c.M);
  k.dtMkc = (T)(c.dt * c.M * c.kappa_c);
  k.dtL = (T)(c.dt * c.L);
  k.dtLke = (T)(c.dt * c.L * c.kappa);
  k.xi = (T)c.xi;
  k.S = ref ? T(0) : (T)c.S;
  k.celastic = ref ? 0 : 1;
  k.noise = (T)c.noise;
  k.cbar = (T)s->cbar;
  k.kfx = (T)kf_of(c, s->nx, w.dx); k.kfy = (T)kf_of(c, s->ny, w.dy); k.kfz = (T)kf_of(c, s->nz, w.dz);
  k.d = BoxDim{s->nx, s->ny, s->nz};
  k.inv_n3 = (T)(1.0 / (double)nvox(s));
  k.V = s->V;
  k.step = (uint32_t)s->step;
  k.seed = c.seed;
  return k;
}

// R_p = IFFT[sum_q B_pq FFT(eta_q^2)] / N from the V fields at eta into R as packed pairs
// (R_2j + i R_2j+1 in complex field j). Every field on the way is real, so the squares and the
// R_p share transforms two by two: ceil(V / 2) forward and ceil(V / 2) inverse transforms.
template <typename T> int elastic_R(gsk_sim* s, const T* eta) {
  BoxWS& w = *s->bx;
  const long N3 = (long)nvox(s);
  const int P = (s->V + 1) / 2;
  C2<T>* S = (C2<T>*)w.S;
  C2<T>* R = (C2<T>*)w.R;
  PlainKT<T> k = mart_k<T>(s);
  LAUNCH(p_square_pair<T>, grid_1d(N3 / 4), 256, 0, s, eta, S, N3, s->V);
  CHK(c2c(s, S, true, P));
  switch (s->V) {
#define GSK_BOX_CONTRACT(V) \
  case V: LAUNCH((p_contract_pair<T, V>), grid_1d(N3), 256, 0, s, S, R, (const T*)w.B, N3, k); break;
    GSK_BOX_CONTRACT(1) GSK_BOX_CONTRACT(2) GSK_BOX_CONTRACT(3) GSK_BOX_CONTRACT(4)
    GSK_BOX_CONTRACT(5) GSK_BOX_CONTRACT(6) GSK_BOX_CONTRACT(7) GSK_BOX_CONTRACT(8)
#undef GSK_BOX_CONTRACT
    default: return fail(GSK_ERR_ARG, "nvar out of range");
  }
  return c2c(s, R, false, P);
}

// Martensite, spectral rule (the plain path's step on a box). Each force shares its transform with
// its own field (F_p + i eta_p: V forward transforms) and the new fields share theirs in pairs
// (ceil(V / 2) inverse transforms).
template <typename T> int step_mart(gsk_sim* s) {
  BoxWS& w = *s->bx;
  const long N3 = (long)nvox(s);
  T* fld = (T*)w.fld;
  C2<T>* E = (C2<T>*)w.E;
  C2<T>* Fh = (C2<T>*)w.Fh;
  const PlainKT<T> k = mart_k<T>(s);
  CHK(elastic_R<T>(s, fld));
  LAUNCH(p_force_pair<T>, grid_1d(N3 / 4), 256, 0, s, fld, (const C2<T>*)w.R, Fh, N3, k);
  CHK(c2c(s, Fh, true, s->V));
  LAUNCH(p_update_pair<T>, grid_1d(N3), 256, 0, s, Fh, E, k);
  CHK(c2c(s, E, false, (s->V + 1) / 2));
  LAUNCH(p_extract_pair<T>, grid_1d(N3 / 4), 256, 0, s, E, fld, N3, s->V);
  return GSK_OK;
}

// Precipitate, spectral rule (c-sourced elasticity, scalar B) or reference rule (eta^2 source with
// the V x V tensor, noise on c, no elasticity in the c update)
template <typename T> int step_prec(gsk_sim* s) {
  BoxWS& w = *s->bx;
  const long N3 = (long)nvox(s);
  const bool ref = s->cfg.rule == GSK_RULE_REFERENCE;
  T* fld = (T*)w.fld;
  C2<T>* E = (C2<T>*)w.E;
  C2<T>* Fh = (C2<T>*)w.Fh;
  const PlainPKT<T> k = prec_k<T>(s);
  if (ref) CHK(elastic_R<T>(s, fld + N3));
  // forces and fields share transforms (F_f + i field_f), the new fields share theirs in pairs
  LAUNCH(p_prec_local<T>, grid_1d(N3 / 4), 256, 0, s, fld, Fh, N3, k, ref ? (const C2<T>*)w.R : nullptr);
  CHK(c2c(s, Fh, true, s->F));
  LAUNCH(p_prec_update_pair<T>, grid_1d(N3), 256, 0, s, Fh, E, ref ? nullptr : (const T*)w.B, k);
  CHK(c2c(s, E, false, (s->F + 1) / 2));
  LAUNCH(p_extract_pair<T>, grid_1d(N3 / 4), 256, 0, s, E, fld, N3, s->F);
  return GSK_OK;
}

// Martensite, reference rule (FP32, explicit Euler, src/dynamics_mart.cu:181-417)
int step_mart_ref(gsk_sim* s) {
  BoxWS& w = *s->bx;
  const gsk_config& c = s->cfg;
  const long N3 = (long)nvox(s);
  float* fld = (float*)w.fld;
  CHK(elastic_R<float>(s, fld));
  LAUNCH(p_extract_pair<float>, grid_1d(N3 / 4), 256, 0, s, (const float2*)w.R, s->rbuf, N3, s->V);
  RefMartK k{};
  k.a1 = (float)(0.02 * (c.temperature - c.T0));
  k.a2 = (float)c.B;
  k.a3 = (float)c.C;
  k.sc = (float)std::sqrt(std::fabs((0.02 * (c.temperature - c.T0)) / (c.B - c.C)));
  k.wG = (float)c.w_grad; k.wC = (float)c.w_chem; k.wE = (float)c.w_el; k.wD = (float)c.w_dis;
  k.wN = (float)c.noise; k.dt = (float)c.dt;
  k.gx = (float)(1.0 / (2.0 * w.dx) / 3.0);
  k.gy = (float)(1.0 / (2.0 * w.dy) / 3.0);
  k.gz = (float)(1.0 / (2.0 * w.dz) / 3.0);
  k.V = s->V;
  k.chem_form = c.chem_form;
  k.step = (uint32_t)s->step;
  k.stepp = nullptr;
  k.seed = c.seed;
  CHK(launch_ref_mart(s, fld, s->rbuf, s->nx, s->ny, s->nz, k));
  // the new eta is in rbuf: the two buffers trade places instead of copying V fields back
  float* next = s->rbuf;
  s->rbuf = fld;
  w.fld = next;
  return GSK_OK;
}

// pass shapes of one axis: z rows per tile (L) or chunk width (W) of the strided passes
void shapes(gsk_sim* s) {
  BoxWS& w = *s->bx;
  const size_t c2 = w.f64 ? sizeof(double2) : sizeof(float2);
  const long rows = (long)s->nx * s->ny;
  const int nz = s->nz, rz = rmin_of(nz);
  int L = 1;
  while (L * 2 <= rows && rows % (L * 2) == 0 && (L * 2) * nz / rz <= 256 &&
         (double)(nz + (L * 2) * (nz + 1)) * c2 <= 96 * 1024)
    L *= 2;
  w.L = L;
  w.tz = round64(std::max(L * nz / rz, 64));
  w.ldz = (size_t)(nz + L * (nz + 1)) * c2;
  for (int axis = 0; axis < 2; ++axis) {
    const int N = axis == 0 ? s->nx : s->ny, r = rmin_of(N);
    int W = 16;
    while (W > 1 && (W * N / r > 1024 || (double)(N + N * W) * c2 > 96 * 1024 || nz % W)) W /= 2;
    w.W[axis] = W;
    const int th = round64(std::max(W * N / r, 64));
    const size_t lds = (size_t)(N + N * W) * c2;
    if (axis == 0) { w.tx = th; w.ldx = lds; }
    else { w.ty = th; w.ldy = lds; }
  }
}

int attrs_of(int n) {
  switch (n) {
    case 16: return BoxN<16>::attrs();
    case 32: return BoxN<32>::attrs();
    case 64: return BoxN<64>::attrs();
    case 96: return BoxN<96>::attrs();
    case 128: return BoxN<128>::attrs();
    case 192: return BoxN<192>::attrs();
    case 256: return BoxN<256>::attrs();
    case 512: return BoxN<512>::attrs();
  }
  return fail(GSK_ERR_UNSUPPORTED, "unsupported axis length " + std::to_string(n));
}

template <typename T> int energy_t(gsk_sim* s, double* gr, double* el) {
  BoxWS& w = *s->bx;
  const long N3 = (long)nvox(s);
  const gsk_config& c = s->cfg;
  T* fld = (T*)w.fld;
  C2<T>* E = (C2<T>*)w.E;
  const int nb = 256;
  double h[3] = {0.0, 0.0, 0.0};
  if (c.physics == GSK_PHYS_MART) {
    const long VN = (long)s->V * N3;
    const PlainKT<T> k = mart_k<T>(s);
    C2<T>* S = (C2<T>*)w.S;
    LAUNCH(p_square<T>, grid_1d(VN), 256, 0, s, fld, S, VN);
    CHK(c2c(s, S, true, s->V));
    LAUNCH(p_load<T>, grid_1d(VN), 256, 0, s, fld, E, VN);
    CHK(c2c(s, E, true, s->V));
    LAUNCH(r_plain_kspace<T>, nb, 256, 0, s, E, S, (const T*)w.B, k, s->part);
    LAUNCH(r_finish<>, 1, 1024, 0, s, s->part, nb, 2, s->rout);
    HIPCHK(hipMemcpyAsync(h, s->rout, 2 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    CHK(sync(s));
    *gr = 0.5 * c.kappa * h[0] / (double)N3;
    *el = 0.5 * c.w_el * h[1] / (double)N3;
  } else {
    const long FN = (long)s->F * N3;
    const PlainPKT<T> k = prec_k<T>(s);
    LAUNCH(p_load<T>, grid_1d(FN), 256, 0, s, fld, E, FN);
    CHK(c2c(s, E, true, s->F));
    LAUNCH(r_plain_kspace_prec<T>, nb, 256, 0, s, E, (const T*)w.B, k, s->part);
    LAUNCH(r_finish<>, 1, 1024, 0, s, s->part, nb, 3, s->rout);
    HIPCHK(hipMemcpyAsync(h, s->rout, 3 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    // the stored concentration is c - cbar: its k = 0 mode f0 stands for c0 = N cbar + f0 in the
    // elastic sum, whose k = 0 weight is B(0) = C eps eps
    C2<T> e0;
    HIPCHK(hipMemcpyAsync(&e0, E, sizeof(e0), hipMemcpyDeviceToHost, s->stream));
    CHK(sync(s));
    const double f0 = (double)e0.x, c0 = (double)N3 * s->cbar + f0;
    *gr = 0.5 * (c.kappa_c * h[0] + c.kappa * h[1]) / (double)N3;
    *el = 0.5 * c.xi * (h[2] - s->cee0 * f0 * f0 + s->cee0 * c0 * c0) / (double)N3;
  }
  return GSK_OK;
}

}  // namespace

bool box_grid(const gsk_config& c) {
  const int ny = c.ny > 0 ? c.ny : c.n, nz = c.nz > 0 ? c.nz : c.n;
  const double dy = c.dy > 0 ? c.dy : c.dx, dz = c.dz > 0 ? c.dz : c.dx;
  return !(ny == c.n && nz == c.n && dy == c.dx && dz == c.dx);
}

int box_create(gsk_sim* s) {
  const gsk_config& c = s->cfg;
  if (c.physics == GSK_PHYS_MART && c.rule == GSK_RULE_SPECTRAL && s->V > 8)
    return fail(GSK_ERR_ARG, "nvar out of range");
  for (int n : {s->nx, s->ny, s->nz}) CHK(attrs_of(n));
  BoxWS* w = new BoxWS();
  s->bx = w;
  w->f64 = c.path == GSK_PATH_PLAIN64;
  w->dx = c.dx;
  w->dy = c.dy > 0 ? c.dy : c.dx;
  w->dz = c.dz > 0 ? c.dz : c.dx;
  shapes(s);
  const size_t N3 = nvox(s), r = w->f64 ? sizeof(double) : sizeof(float), cx = 2 * r;
  auto alloc = [&](void*& p, size_t bytes) -> int {
    if (hipMalloc(&p, bytes) != hipSuccess)
      return fail(GSK_ERR_HIP, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    if (hipMemsetAsync(p, 0, bytes, s->stream) != hipSuccess) return fail(GSK_ERR_HIP, "hipMemsetAsync failed");
    return GSK_OK;
  };
  const bool vv = c.physics == GSK_PHYS_MART || c.rule == GSK_RULE_REFERENCE;  // V x V tensor on eta^2
  CHK(alloc(w->fld, (size_t)s->F * N3 * r));
  CHK(alloc(w->E, (size_t)s->F * N3 * cx));
  CHK(alloc(w->Fh, (size_t)s->F * N3 * cx));
  if (vv) {
    CHK(alloc(w->S, (size_t)s->V * N3 * cx));
    CHK(alloc(w->R, (size_t)s->V * N3 * cx));
  }
  CHK(alloc(w->B, (vv ? (size_t)s->V * s->V : 1) * N3 * r));
  const int ns[3] = {s->nx, s->ny, s->nz};
  for (int a = 0; a < 3; ++a) {
    const int n = ns[a];
    if (w->f64) {
      const auto t = twiddles<double2>(n, n);
      CHK(alloc(w->tw[a], t.size() * sizeof(double2)));
      HIPCHK(hipMemcpy(w->tw[a], t.data(), t.size() * sizeof(double2), hipMemcpyHostToDevice));
    } else {
      const auto t = twiddles<float2>(n, n);
      CHK(alloc(w->tw[a], t.size() * sizeof(float2)));
      HIPCHK(hipMemcpy(w->tw[a], t.data(), t.size() * sizeof(float2), hipMemcpyHostToDevice));
    }
  }
  if (c.rule == GSK_RULE_REFERENCE && c.physics == GSK_PHYS_MART) {
    if (hipMalloc((void**)&s->rbuf, (size_t)s->V * N3 * sizeof(float)) != hipSuccess ||
        hipMalloc((void**)&s->fstat, (size_t)kRefBlocks * 4 * sizeof(double)) != hipSuccess)
      return fail(GSK_ERR_HIP, "hipMalloc failed");
    HIPCHK(hipMemsetAsync(s->fstat, 0, (size_t)kRefBlocks * 4 * sizeof(double), s->stream));
  }
  return GSK_OK;
}

void box_free(BoxWS* w) {
  if (!w) return;
  for (void* p : {w->fld, w->S, w->R, w->E, w->Fh, w->B, w->tw[0], w->tw[1], w->tw[2]})
    if (p) (void)hipFree(p);
  delete w;
}

int box_build_btensor(gsk_sim* s, const BArgs& a) {
  BoxWS& w = *s->bx;
  const gsk_config& c = s->cfg;
  const int npq = a.V1 * a.V2;
  const double kx = kf_of(c, s->nx, w.dx), ky = kf_of(c, s->ny, w.dy), kz = kf_of(c, s->nz, w.dz);
  const long N3 = (long)nvox(s);
  if (w.f64)
    hipLaunchKernelGGL(k_btensor_box<double>, dim3(grid_1d(N3)), dim3(256), 0, s->stream, (double*)w.B, s->nx, s->ny,
                       s->nz, kx, ky, kz, a, npq);
  else
    hipLaunchKernelGGL(k_btensor_box<float>, dim3(grid_1d(N3)), dim3(256), 0, s->stream, (float*)w.B, s->nx, s->ny,
                       s->nz, kx, ky, kz, a, npq);
  HIPCHK(hipGetLastError());
  return GSK_OK;
}

// The FP32 path keeps the concentration as cbar + an FP32 fluctuation, cbar its FP64 mean rounded to
// binary32 (so c - cbar is exact in FP32 near the mean): FP32 rounding then acts on the fluctuation,
// not on the mean. FP64 fields are stored as they are (cbar = 0).
int box_set_field(gsk_sim* s, int off, size_t count, const double* h) {
  BoxWS& w = *s->bx;
  const size_t N3 = nvox(s);
  CHK(sync(s));  // earlier work on the handle's stream may still read the field
  const bool conc = s->cfg.physics == GSK_PHYS_PREC && off == 0;
  if (w.f64) {
    HIPCHK(hipMemcpy((double*)w.fld + (size_t)off * N3, h, count * sizeof(double), hipMemcpyHostToDevice));
  } else {
    double cbar = 0.0;
    if (conc) {
      double sum = 0.0, comp = 0.0;  // compensated sum of the concentration
      for (size_t i = 0; i < N3; ++i) {
        const double y = h[i] - comp, t = sum + y;
        comp = (t - sum) - y;
        sum = t;
      }
      cbar = (double)(float)(sum / (double)N3);
      s->cbar = cbar;
    }
    std::vector<float> f(count);
    for (size_t i = 0; i < count; ++i) f[i] = (float)(i < N3 ? h[i] - cbar : h[i]);
    HIPCHK(hipMemcpy((float*)w.fld + (size_t)off * N3, f.data(), count * sizeof(float), hipMemcpyHostToDevice));
  }
  return GSK_OK;
}

int box_get_field(gsk_sim* s, int off, size_t count, double* h) {
  BoxWS& w = *s->bx;
  const size_t N3 = nvox(s);
  CHK(sync(s));
  if (w.f64) {
    HIPCHK(hipMemcpy(h, (const double*)w.fld + (size_t)off * N3, count * sizeof(double), hipMemcpyDeviceToHost));
  } else {
    std::vector<float> f(count);
    HIPCHK(hipMemcpy(f.data(), (const float*)w.fld + (size_t)off * N3, count * sizeof(float), hipMemcpyDeviceToHost));
    const double cbar = (s->cfg.physics == GSK_PHYS_PREC && off == 0) ? s->cbar : 0.0;
    for (size_t i = 0; i < count; ++i) h[i] = (i < N3 ? cbar : 0.0) + (double)f[i];
  }
  return GSK_OK;
}

int box_step(gsk_sim* s) {
  const gsk_config& c = s->cfg;
  const bool f64 = s->bx->f64;
  if (c.physics == GSK_PHYS_MART) {
    if (c.rule == GSK_RULE_REFERENCE) return step_mart_ref(s);
    return f64 ? step_mart<double>(s) : step_mart<float>(s);
  }
  return f64 ? step_prec<double>(s) : step_prec<float>(s);
}

int box_energy(gsk_sim* s, double* gr, double* el) {
  return s->bx->f64 ? energy_t<double>(s, gr, el) : energy_t<float>(s, gr, el);
}
