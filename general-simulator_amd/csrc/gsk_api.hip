// C ABI of the MI355X phase-field step library. See include/gsk.h.
#include "gsk.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <type_traits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "kernels_aux.h"
#include "kernels_adjoint.h"
#include "kernels_ref.h"
#include "kernels_plane.h"

using namespace gsk;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                                  \
  do {                                                                                                \
    hipError_t e_ = (expr);                                                                           \
    if (e_ != hipSuccess)                                                                             \
      return fail(GSK_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_) + " (" + __FILE__ + \
                                   ":" + std::to_string(__LINE__) + ")");                             \
  } while (0)

#define CHK(expr)                  \
  do {                             \
    int rc_ = (expr);              \
    if (rc_ != GSK_OK) return rc_; \
  } while (0)

constexpr double kPi = 3.14159265358979323846;

bool supported_n(int n) {
  switch (n) {
    case 16: case 32: case 64: case 96: case 128: case 192: case 256: case 512: return true;
    default: return false;
  }
}

int rmin_of(int N) {
  switch (N) {
    case 8: return PlanOf<8>::rmin;
    case 16: return PlanOf<16>::rmin;
    case 32: return PlanOf<32>::rmin;
    case 48: return PlanOf<48>::rmin;
    case 64: return PlanOf<64>::rmin;
    case 96: return PlanOf<96>::rmin;
    case 128: return PlanOf<128>::rmin;
    case 192: return PlanOf<192>::rmin;
    case 256: return PlanOf<256>::rmin;
    case 512: return PlanOf<512>::rmin;
  }
  return 1;
}

int round64(int t) { return ((t + 63) / 64) * 64; }

// Launch shapes of the strided passes: GSK_Y_PERFIELD gives every field its own y-pass
// work groups; GSK_XUPD_G is the largest field group of one x-update launch.
#ifndef GSK_Y_PERFIELD
#define GSK_Y_PERFIELD 1
#endif
#ifndef GSK_XUPD_G
#define GSK_XUPD_G 0
#endif

#ifndef GSK_FREV
#define GSK_FREV 1
#endif
#ifndef GSK_TREV
#define GSK_TREV 0
#endif

// fields in the next martensite x-update launch when `left` fields remain
int xupd_group(int left) {
  if (GSK_XUPD_G == 0) return left;  // a single launch covers every field
  return (GSK_XUPD_G >= 3 && left >= 3) ? 3 : ((GSK_XUPD_G >= 2 && left >= 2) ? 2 : 1);
}

// block size of the z kernels for half-length M
int zthreads_host(int M) {
  switch (M) {
    case 8: return zthreads<8>();
    case 16: return zthreads<16>();
    case 32: return zthreads<32>();
    case 48: return zthreads<48>();
    case 64: return zthreads<64>();
    case 96: return zthreads<96>();
    case 128: return zthreads<128>();
    case 256: return zthreads<256>();
  }
  return 0;
}

template <typename T2>
std::vector<T2> twiddles(int N, int denom) {
  std::vector<T2> t(N);
  for (int m = 0; m < N; ++m) {
    double a = -2.0 * kPi * m / denom;
    t[m].x = (decltype(t[m].x))std::cos(a);
    t[m].y = (decltype(t[m].x))std::sin(a);
  }
  return t;
}

struct Slot {
  const char* name;
  double bytes;  // algorithmic HBM bytes per step
};

}  // namespace

struct AdjWS;  // adjoint workspace (checkpoints and buffers), defined with the adjoint entry points

struct gsk_sim {
  gsk_config cfg{};
  int n = 0, V = 0, F = 0, dev = 0;
  Geo g{};
  int zthreads = 0, xthreads = 0;
  size_t lds_z = 0, lds_y = 0, lds_xc = 0;
  hipStream_t stream = nullptr;
  // twiddles
  float2 *twN = nullptr, *twM = nullptr, *rtn = nullptr;
  double2* twN64 = nullptr;
  // fused FP32 buffers
  float* fld = nullptr;
  float2* kbuf = nullptr;
  float2* state = nullptr;
  float* Bt = nullptr;
  // plain FP64 buffers
  double* eta64 = nullptr;
  double2 *S64 = nullptr, *R64 = nullptr, *E64 = nullptr, *F64 = nullptr;
  double* B64 = nullptr;
  int pL = 0, pW = 0, pzthreads = 0, pxthreads = 0;
  size_t lds_pz = 0, lds_px = 0;
  // reductions
  double* part = nullptr;
  double* rout = nullptr;
  uint32_t* dstep = nullptr;  // fused path: device copy of the step index (noise counter)
  unsigned gen = 0;           // bumped whenever a value baked into kernel arguments changes
  BArgs bargs{};              // inputs of the last B tensor build (the adjoint rebuilds B in its own layout)
  bool haveBargs = false;
  AdjWS* adj = nullptr;
  // reference rules
  float* rbuf = nullptr;      // V real fields: elastic R, then the new eta (martensite)
  float* defect = nullptr;    // n^3 blocking field (martensite reference rule), optional
  float* Kd = nullptr;        // V*n^3 dislocation force constants, optional
  double eps[72] = {};        // eigenstrains of the last B build
  bool haveEps = false;
  bool sharedB = false;       // Bt belongs to an ensemble, not to this handle
  int rblocks = 512;
  // state
  double cbar = 0.0;
  double c_k0 = 0.0;  // FP64 sum of the stored concentration fluctuation (its k = 0 mode)
  double cee0 = 0.0;  // B(k = 0) of the c-sourced tensor
  bool haveB = false, dirty = true;
  // martensite fused step: the k-space buffer holds the z-transformed eta^2 of the current
  // fields (left there by the last step's fused inverse/square pass); cleared by anything that
  // writes the fields or uses the buffer otherwise
  bool sq_ready = false;
  int order = 0;  // visiting order of the last per-field launch
  long long step = 0;
  // profiling
  bool prof = false;
  std::vector<hipEvent_t> ev;
  int evi = 0;
  std::vector<Slot> slots;
};

namespace {

// Geometry for the next per-field launch: flips the field / tile visiting order each time.
Geo geo_next(gsk_sim* s) {
  Geo g = s->g;
  s->order ^= 1;
  g.frev = GSK_FREV ? s->order : 0;
  g.trev = GSK_TREV ? s->order : 0;
  return g;
}

// Plane-resident passes of the spectral martensite step (n = 32, 64, 128 and 1..4 variants):
// k_plane_mid replaces y inverse -> middle z pass -> y forward, k_plane_sq the y inverse -> last
// z pass -> next step's y forward, so the k-space buffer carries the y,z-transformed eta^2
// between steps (z-transformed only at the other sizes). GSK_PLANE=0 in the environment keeps
// the chain of y and z kernels.
bool plane_on(const gsk_sim* s) {
  static const bool off = [] {
    const char* e = std::getenv("GSK_PLANE");
    return e && e[0] == '0';
  }();
  return !off && (s->n == 32 || s->n == 64 || s->n == 128) && s->F >= 1 && s->F <= 4 &&
         s->cfg.path == GSK_PATH_FUSED32 && s->cfg.physics == GSK_PHYS_MART && s->cfg.rule == GSK_RULE_SPECTRAL;
}

size_t nvox(const gsk_sim* s) { return (size_t)s->n * s->n * s->n; }
size_t hs_count(const gsk_sim* s) { return (size_t)s->n * s->n * (s->n / 2 + 1); }
// State element of field 0 at k = 0: first main x tile (after the n/W Nyquist tiles), kx = 0, column 0.
size_t k0_index(const gsk_sim* s) { return (size_t)(s->n / s->g.W) * s->F * s->n * s->g.W; }

int mark(gsk_sim* s) {
  if (s->prof) {
    if (s->evi >= (int)s->ev.size()) return fail(GSK_ERR, "profiling event overflow");
    HIPCHK(hipEventRecord(s->ev[s->evi++], s->stream));
  }
  return GSK_OK;
}

#define LAUNCH(kern, grid, threads, lds, s, ...)                                              \
  do {                                                                                      \
    CHK(mark(s));                                                                           \
    hipLaunchKernelGGL(kern, dim3(grid), dim3(threads), lds, (s)->stream, __VA_ARGS__);     \
    HIPCHK(hipGetLastError());                                                              \
  } while (0)

MartK mart_consts(const gsk_sim* s) {
  const gsk_config& c = s->cfg;
  MartK k{};
  k.A = (float)c.A; k.B = (float)c.B; k.C = (float)c.C;
  k.two_wel = (float)(2.0 * c.w_el);
  k.noise = (float)c.noise;
  k.dtL = (float)(c.dt * c.L);
  k.dtLk = (float)(c.dt * c.L * c.kappa);
  k.kfac = (float)(c.kconv == GSK_K_REFERENCE ? 2.0 * 3.14 * c.dx / s->n : 2.0 * kPi / (s->n * c.dx));
  k.inv_n3 = (float)(1.0 / ((double)s->n * s->n * s->n));
  k.F = s->V;
  k.npq = s->V * (s->V + 1) / 2;
  k.step = (uint32_t)s->step;
  k.stepp = s->dstep;
  k.seed = c.seed;
  return k;
}

PrecK prec_consts(const gsk_sim* s) {
  const gsk_config& c = s->cfg;
  PrecK k{};
  k.A1 = (float)c.a[1]; k.A2 = (float)c.a[2]; k.A3 = (float)c.a[3]; k.A4 = (float)c.a[4];
  k.A5 = (float)c.a[5]; k.A6 = (float)c.a[6]; k.A7 = (float)c.a[7];
  k.c1 = (float)c.c1; k.c2 = (float)c.c2;
  k.cbar = (float)s->cbar;
  k.k0 = (float)s->c_k0;
  k.noise = (float)c.noise;
  k.dtM = (float)(c.dt * c.M);
  k.dtMkc = (float)(c.dt * c.M * c.kappa_c);
  k.dtL = (float)(c.dt * c.L);
  k.dtLke = (float)(c.dt * c.L * c.kappa);
  k.xi = (float)c.xi;
  k.S = c.rule == GSK_RULE_REFERENCE ? 0.f : (float)c.S;
  k.celastic = c.rule == GSK_RULE_REFERENCE ? 0 : 1;
  k.kfac = (float)(c.kconv == GSK_K_REFERENCE ? 2.0 * 3.14 * c.dx / s->n : 2.0 * kPi / (s->n * c.dx));
  k.inv_n3 = (float)(1.0 / ((double)s->n * s->n * s->n));
  k.V = s->V;
  k.step = (uint32_t)s->step;
  k.stepp = s->dstep;
  k.seed = c.seed;
  return k;
}

PlainK plain_consts(const gsk_sim* s) {
  const gsk_config& c = s->cfg;
  PlainK k{};
  k.A = c.A; k.B = c.B; k.C = c.C;
  k.two_wel = 2.0 * c.w_el;
  k.noise = c.noise;
  k.dtL = c.dt * c.L;
  k.dtLk = c.dt * c.L * c.kappa;
  k.kfac = (c.kconv == GSK_K_REFERENCE ? 2.0 * 3.14 * c.dx / s->n : 2.0 * kPi / (s->n * c.dx));
  k.inv_n3 = 1.0 / ((double)s->n * s->n * s->n);
  k.V = s->V;
  k.step = (uint32_t)s->step;
  k.seed = c.seed;
  return k;
}

int grid_1d(long work, int threads = 256) {
  long b = (work + threads - 1) / threads;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (int)b;
}

// Allow a kernel the device's whole LDS for dynamic shared memory (minus its static part).
int lds_limit(const void* f) {
  int dev = 0, optin = 0, maxs = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&maxs, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev) == hipSuccess && optin > maxs)
    maxs = optin;
  hipFuncAttributes a;
  HIPCHK(hipFuncGetAttributes(&a, f));
  HIPCHK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, maxs - (int)a.sharedSizeBytes));
  return GSK_OK;
}

// Steps whose last z pass leaves the first k-space stage of the next step in the k-space buffer
// (the spectral rules on the fused path; the precipitate with 2..5 fields).
bool prec_carries(const gsk_sim* s) {
  return s->cfg.path == GSK_PATH_FUSED32 && s->cfg.physics == GSK_PHYS_PREC && s->cfg.rule == GSK_RULE_SPECTRAL &&
         s->F >= 2 && s->F <= 5;
}
bool carries(const gsk_sim* s) {
  return prec_carries(s) || (s->cfg.path == GSK_PATH_FUSED32 && s->cfg.physics == GSK_PHYS_MART &&
                             s->cfg.rule == GSK_RULE_SPECTRAL);
}

// ------------------------------------------------------------------ per-size implementation
template <int N>
struct Impl {
  static constexpr int M = N / 2;

  static int set_attrs() {
    CHK(lds_limit((const void*)k_zfwd<M, Z_COPY>));
    CHK(lds_limit((const void*)k_zfwd<M, Z_SQUARE>));
    CHK(lds_limit((const void*)k_zfwd_local<M>));
    CHK(lds_limit((const void*)k_zfwd_local_r<M, 2>));
    CHK(lds_limit((const void*)k_zfwd_local_r<M, 3>));
    CHK(lds_limit((const void*)k_zfwd_local_r<M, 4>));
    CHK(lds_limit((const void*)k_zfwd_local_r<M, 5>));
    CHK(lds_limit((const void*)k_zmid_mart<M, 0>));
    CHK(lds_limit((const void*)k_zmid_mart<M, 1>));
    CHK(lds_limit((const void*)k_zmid_mart<M, 2>));
    CHK(lds_limit((const void*)k_zmid_mart<M, 3>));
    CHK(lds_limit((const void*)k_zmid_mart<M, 4>));
    CHK(lds_limit((const void*)k_zinv<M>));
    CHK(lds_limit((const void*)k_zinv_sq<M>));
    CHK(lds_limit((const void*)k_zinv_local_r<M, 2>));
    CHK(lds_limit((const void*)k_zinv_local_r<M, 3>));
    CHK(lds_limit((const void*)k_zinv_local_r<M, 4>));
    CHK(lds_limit((const void*)k_zinv_local_r<M, 5>));
    CHK(lds_limit((const void*)k_ypass<N, true>));
    CHK(lds_limit((const void*)k_ypass<N, false>));
    CHK(lds_limit((const void*)k_xpass<N, true>));
    CHK(lds_limit((const void*)k_xpass<N, false>));
    CHK(lds_limit((const void*)k_xcontract<N>));
    CHK(lds_limit((const void*)k_xcontract_r<N, 1>));
    CHK(lds_limit((const void*)k_xcontract_r<N, 2>));
    CHK(lds_limit((const void*)k_xcontract_r<N, 3>));
    CHK(lds_limit((const void*)k_xcontract_r<N, 4>));
    CHK(lds_limit((const void*)k_xupdate_mart<N, 1, 0>));
    CHK(lds_limit((const void*)k_xupdate_mart<N, 2, 0>));
    CHK(lds_limit((const void*)k_xupdate_mart<N, 1, 1>));
    CHK(lds_limit((const void*)k_xupdate_mart<N, 2, 1>));
    CHK(lds_limit((const void*)k_xupdate_mart<N, 3, 0>));
    CHK(lds_limit((const void*)k_xupdate_mart<N, 3, 1>));
    CHK(lds_limit((const void*)k_xupdate_prec<N, 1, 0>));
    CHK(lds_limit((const void*)k_xupdate_prec<N, 2, 0>));
    CHK(lds_limit((const void*)k_xupdate_prec<N, 1, 1>));
    CHK(lds_limit((const void*)k_xupdate_prec<N, 2, 1>));
    CHK(lds_limit((const void*)r_xenergy_mart<N>));
    CHK(lds_limit((const void*)r_xenergy_mart_r<N, 1>));
    CHK(lds_limit((const void*)r_xenergy_mart_r<N, 2>));
    CHK(lds_limit((const void*)r_xenergy_mart_r<N, 3>));
    CHK(lds_limit((const void*)r_xenergy_mart_r<N, 4>));
    CHK(lds_limit((const void*)p_zrows<N, true>));
    CHK(lds_limit((const void*)p_zrows<N, false>));
    CHK(lds_limit((const void*)p_strided<N, true>));
    CHK(lds_limit((const void*)p_strided<N, false>));
    if constexpr (plane_size<N>()) {
      CHK(lds_limit((const void*)k_plane_sq<N>));
      CHK(lds_limit((const void*)k_plane_sq0<N>));
      CHK(lds_limit((const void*)k_plane_mid<N, 1>));
      CHK(lds_limit((const void*)k_plane_mid<N, 2>));
      CHK(lds_limit((const void*)k_plane_mid<N, 3>));
      CHK(lds_limit((const void*)k_plane_mid<N, 4>));
    }
    return GSK_OK;
  }

  // y inverse of eta, last z pass (eta written, squared) and y forward in one launch, one x-plane
  // of one field per work group (plane_on); adv: the step counter the launch advances
  static int plane_sq(gsk_sim* s, uint32_t* adv) {
    if constexpr (plane_size<N>()) {
      LAUNCH((k_plane_sq<N>), dim3(N, s->F), plane_threads<N>(), plane_lds<N>(), s, s->kbuf, s->fld, geo_next(s), s->twN,
             s->twM, s->rtn, adv, (const MemberPtrs*)nullptr);
      return GSK_OK;
    } else {
      (void)s; (void)adv;
      return fail(GSK_ERR_UNSUPPORTED, "the plane pass takes n = 32, 64, 128");
    }
  }

  // y inverse of R, middle z pass and y forward of F in one launch, one x-plane of one variant
  // per work group (plane_on)
  static int plane_mid(gsk_sim* s, const MartK& mk) {
    if constexpr (plane_size<N>()) {
      const dim3 grid(N, s->F);
      const Geo g = geo_next(s);
      constexpr int T = plane_threads<N>();
      constexpr size_t lds = plane_lds<N>();
      switch (s->F) {
        case 1: LAUNCH((k_plane_mid<N, 1>), grid, T, lds, s, s->kbuf, s->fld, g, mk, s->twN, s->twM, s->rtn, (const MemberPtrs*)nullptr); break;
        case 2: LAUNCH((k_plane_mid<N, 2>), grid, T, lds, s, s->kbuf, s->fld, g, mk, s->twN, s->twM, s->rtn, (const MemberPtrs*)nullptr); break;
        case 3: LAUNCH((k_plane_mid<N, 3>), grid, T, lds, s, s->kbuf, s->fld, g, mk, s->twN, s->twM, s->rtn, (const MemberPtrs*)nullptr); break;
        case 4: LAUNCH((k_plane_mid<N, 4>), grid, T, lds, s, s->kbuf, s->fld, g, mk, s->twN, s->twM, s->rtn, (const MemberPtrs*)nullptr); break;
        default: return fail(GSK_ERR_UNSUPPORTED, "the plane pass takes 1..4 variants");
      }
      return GSK_OK;
    } else {
      (void)s; (void)mk;
      return fail(GSK_ERR_UNSUPPORTED, "the plane pass takes n = 32, 64, 128");
    }
  }

  // ---- fused building blocks
  // LDS of a z kernel holding nf field tiles
  static size_t zlds(const gsk_sim* s, int nf) { return (2 * M + (size_t)nf * s->g.L * zpitch(M)) * sizeof(float2); }
  // shape of the middle martensite pass (its own rows per tile)
  static constexpr int LZ = zmid_rows(M), ZMT = zmid_threads<M>();
  static constexpr size_t ZM_LDS = (2 * M + (size_t)LZ * zpitch_l(M, LZ)) * sizeof(float2);
  static int zfwd(gsk_sim* s, int op, const float* in, int F) {
    const int rows = s->n * s->n / s->g.L;
    if (op == Z_COPY) {
      LAUNCH((k_zfwd<M, Z_COPY>), dim3(rows, F), s->zthreads, zlds(s, 1), s, in, s->kbuf, geo_next(s), s->twM, s->rtn);
    } else if (op == Z_SQUARE) {
      LAUNCH((k_zfwd<M, Z_SQUARE>), dim3(rows, F), s->zthreads, zlds(s, 1), s, in, s->kbuf, geo_next(s), s->twM, s->rtn);
    } else {
      PrecK pk = prec_consts(s);
      const bool ref = s->cfg.rule == GSK_RULE_REFERENCE;
      const float* R = ref ? (const float*)s->rbuf : (const float*)nullptr;
      const float el = ref ? (float)(2.0 * s->cfg.xi) : 0.f, cn = ref ? (float)s->cfg.noise : 0.f;
      switch (F) {  // every field of a tile in registers for up to 5 fields
        case 2: LAUNCH((k_zfwd_local_r<M, 2>), rows, s->zthreads, zlds(s, 1), s, in, s->kbuf, s->g, pk, s->twM, s->rtn, R, el, cn); break;
        case 3: LAUNCH((k_zfwd_local_r<M, 3>), rows, s->zthreads, zlds(s, 1), s, in, s->kbuf, s->g, pk, s->twM, s->rtn, R, el, cn); break;
        case 4: LAUNCH((k_zfwd_local_r<M, 4>), rows, s->zthreads, zlds(s, 1), s, in, s->kbuf, s->g, pk, s->twM, s->rtn, R, el, cn); break;
        case 5: LAUNCH((k_zfwd_local_r<M, 5>), rows, s->zthreads, zlds(s, 1), s, in, s->kbuf, s->g, pk, s->twM, s->rtn, R, el, cn); break;
        default: LAUNCH((k_zfwd_local<M>), rows, s->zthreads, zlds(s, F), s, in, s->kbuf, s->g, pk, s->twM, s->rtn, R, el, cn);
      }
    }
    return GSK_OK;
  }
  static int ypass(gsk_sim* s, bool fwd, int F) {
    const dim3 grid(s->g.ntile, GSK_Y_PERFIELD ? F : 1);
    const Geo g = GSK_Y_PERFIELD ? geo_next(s) : s->g;
    if (fwd) LAUNCH((k_ypass<N, true>), grid, s->xthreads, s->lds_y, s, s->kbuf, g, F, s->twN, (const MemberPtrs*)nullptr);
    else LAUNCH((k_ypass<N, false>), grid, s->xthreads, s->lds_y, s, s->kbuf, g, F, s->twN, (const MemberPtrs*)nullptr);
    return GSK_OK;
  }
  static int zinv(gsk_sim* s, float* out, int F, uint32_t* adv = nullptr) {
    const int rows = s->n * s->n / s->g.L;
    LAUNCH((k_zinv<M>), dim3(rows, F), s->zthreads, zlds(s, 1), s, s->kbuf, out, geo_next(s), s->twM, s->rtn, adv);
    return GSK_OK;
  }

  // k-space updates in groups of 2 fields (register-resident x transforms)
  template <int MODE>
  static int xupdate_mart(gsk_sim* s, const MartK& mk) {
    const int gx = s->g.ntile;
    if (GSK_XUPD_G == 0) {  // one field per work group, one launch
      LAUNCH((k_xupdate_mart<N, 1, MODE>), dim3(gx, s->F), s->xthreads, s->lds_y, s, s->kbuf, s->state, geo_next(s), mk,
             s->twN, 0, (const MemberPtrs*)nullptr);
      return GSK_OK;
    }
    for (int f0 = 0; f0 < s->F;) {
      const int G = xupd_group(s->F - f0);
      if (G == 3)
        LAUNCH((k_xupdate_mart<N, 3, MODE>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->state, s->g, mk, s->twN, f0, (const MemberPtrs*)nullptr);
      else if (G == 2)
        LAUNCH((k_xupdate_mart<N, 2, MODE>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->state, s->g, mk, s->twN, f0, (const MemberPtrs*)nullptr);
      else
        LAUNCH((k_xupdate_mart<N, 1, MODE>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->state, s->g, mk, s->twN, f0, (const MemberPtrs*)nullptr);
      f0 += G;
    }
    return GSK_OK;
  }
  template <int MODE>
  static int xupdate_prec(gsk_sim* s, const PrecK& pk) {
    const int gx = s->g.ntile;
    if (GSK_XUPD_G == 0) {  // one field per work group, one launch
      LAUNCH((k_xupdate_prec<N, 1, MODE>), dim3(gx, s->F), s->xthreads, s->lds_y, s, s->kbuf, s->state, s->Bt,
             geo_next(s), pk, s->twN, 0);
      return GSK_OK;
    }
    PrecK pg = pk;
    for (int f0 = 0; f0 < s->F; f0 += 2) {
      if (s->F - f0 >= 2)
        LAUNCH((k_xupdate_prec<N, 2, MODE>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->state, s->Bt, s->g, pg, s->twN,
               f0);
      else
        LAUNCH((k_xupdate_prec<N, 1, MODE>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->state, s->Bt, s->g, pg, s->twN,
               f0);
      pg.adv = nullptr;  // one advance per step
    }
    return GSK_OK;
  }

  static int init_state(gsk_sim* s) {
    if (s->cfg.physics == GSK_PHYS_MART) {
      CHK(zfwd(s, Z_COPY, s->fld, s->F));
      CHK(ypass(s, true, s->F));
      MartK mk = mart_consts(s);
      CHK(xupdate_mart<1>(s, mk));
    } else {
      CHK(zfwd(s, Z_COPY, s->fld, s->F));
      CHK(ypass(s, true, s->F));
      PrecK pk = prec_consts(s);
      CHK(xupdate_prec<1>(s, pk));
    }
    return GSK_OK;
  }

  // z-transformed eta^2 of the current fields into the k-space buffer (y,z-transformed where
  // plane_on): the first kernel of a step, unless the previous step's fused last pass has left
  // it there
  static int square_fwd(gsk_sim* s) {
    if (plane_on(s)) {
      if constexpr (plane_size<N>()) {
        LAUNCH((k_plane_sq0<N>), dim3(N, s->F), plane_threads<N>(), plane_lds<N>(), s, s->fld, s->kbuf, geo_next(s),
               s->twN, s->twM, s->rtn);
      }
    } else {
      CHK(zfwd(s, s->cfg.physics == GSK_PHYS_MART ? Z_SQUARE : Z_PREC_LOCAL, s->fld, s->F));
    }
    s->sq_ready = true;
    return GSK_OK;
  }

  static int step_mart_fused(gsk_sim* s) {
    const int gx = s->g.ntile;
    const int F = s->F;
    MartK mk = mart_consts(s);
    const bool pl = plane_on(s);
    if (!s->sq_ready) CHK(square_fwd(s));
    if (!pl) CHK(ypass(s, true, F));
    switch (F) {  // register-resident contraction for up to 4 variants
      case 1: LAUNCH((k_xcontract_r<N, 1>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->Bt, s->g, mk, s->twN, (const MemberPtrs*)nullptr); break;
      case 2: LAUNCH((k_xcontract_r<N, 2>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->Bt, s->g, mk, s->twN, (const MemberPtrs*)nullptr); break;
      case 3: LAUNCH((k_xcontract_r<N, 3>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->Bt, s->g, mk, s->twN, (const MemberPtrs*)nullptr); break;
      case 4: LAUNCH((k_xcontract_r<N, 4>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->Bt, s->g, mk, s->twN, (const MemberPtrs*)nullptr); break;
      default: LAUNCH((k_xcontract<N>), gx, s->xthreads, s->lds_xc, s, s->kbuf, s->Bt, s->g, mk, s->twN);
    }
    if (pl) {
      CHK(plane_mid(s, mk));
    } else {
      CHK(ypass(s, false, F));
      {
        const int grid = s->n * s->n / LZ;
        const size_t lds = ZM_LDS;
        switch (F) {  // eta prefetched into registers for up to 4 variants
          case 1: LAUNCH((k_zmid_mart<M, 1>), grid, ZMT, lds, s, s->kbuf, s->fld, s->g, mk, s->twM, s->rtn, (const MemberPtrs*)nullptr); break;
          case 2: LAUNCH((k_zmid_mart<M, 2>), grid, ZMT, lds, s, s->kbuf, s->fld, s->g, mk, s->twM, s->rtn, (const MemberPtrs*)nullptr); break;
          case 3: LAUNCH((k_zmid_mart<M, 3>), grid, ZMT, lds, s, s->kbuf, s->fld, s->g, mk, s->twM, s->rtn, (const MemberPtrs*)nullptr); break;
          case 4: LAUNCH((k_zmid_mart<M, 4>), grid, ZMT, lds, s, s->kbuf, s->fld, s->g, mk, s->twM, s->rtn, (const MemberPtrs*)nullptr); break;
          default: LAUNCH((k_zmid_mart<M, 0>), grid, ZMT, lds, s, s->kbuf, s->fld, s->g, mk, s->twM, s->rtn, (const MemberPtrs*)nullptr);
        }
      }
      CHK(ypass(s, true, F));
    }
    CHK(xupdate_mart<0>(s, mk));
    if (pl) {  // eta back to real space, fused with the next step's forward passes of eta^2
      CHK(plane_sq(s, s->dstep));
    } else {  // eta back to real space, fused with the next step's forward z pass of eta^2
      CHK(ypass(s, false, F));
      const int rows = s->n * s->n / s->g.L;
      LAUNCH((k_zinv_sq<M>), dim3(rows, F), s->zthreads, zlds(s, 1), s, s->kbuf, s->fld, geo_next(s), s->twM, s->rtn,
             s->dstep, (const MemberPtrs*)nullptr);
    }
    s->sq_ready = true;
    return GSK_OK;
  }

  // One spectral martensite step of `cnt` members of this grid size in every launch
  // (grid.z = member). The members' buffers and constants come from the device table `tab`;
  // s0, one of them, supplies the geometry, launch shapes and twiddle tables, which members
  // of one size share. Every member must hold its transformed eta^2 (sq_ready; y,z-transformed
  // where plane_on, which holds for all members of one size alike).
  static int step_mart_batched(gsk_sim* s0, int cnt, const MemberPtrs* tab, hipStream_t st) {
    const int F = s0->F, gx = s0->g.ntile, rows = s0->n * s0->n / s0->g.L;
    const MartK mk{};  // replaced by the table entry in every work group
    const size_t lz = zlds(s0, 1);
    float2* nk = nullptr;
    auto y = [&](bool fwd) -> int {
      if (fwd)
        hipLaunchKernelGGL((k_ypass<N, true>), dim3(gx, F, cnt), dim3(s0->xthreads), s0->lds_y, st, nk, geo_next(s0),
                           F, s0->twN, tab);
      else
        hipLaunchKernelGGL((k_ypass<N, false>), dim3(gx, F, cnt), dim3(s0->xthreads), s0->lds_y, st, nk, geo_next(s0),
                           F, s0->twN, tab);
      HIPCHK(hipGetLastError());
      return GSK_OK;
    };
    const bool pl = plane_on(s0);
    if (!pl) CHK(y(true));
    switch (F) {
#define GSK_XC(V)                                                                                                  \
  case V:                                                                                                          \
    hipLaunchKernelGGL((k_xcontract_r<N, V>), dim3(gx, 1, cnt), dim3(s0->xthreads), s0->lds_y, st, nk,            \
                       (const float*)nullptr, s0->g, mk, s0->twN, tab);                                            \
    break;
      GSK_XC(1) GSK_XC(2) GSK_XC(3) GSK_XC(4)
#undef GSK_XC
      default: return fail(GSK_ERR_UNSUPPORTED, "batched ensemble launches take 1..4 variants");
    }
    HIPCHK(hipGetLastError());
    if (pl) {  // the members' own step uses the plane passes: keep them bit-identical
      if constexpr (plane_size<N>()) {
        switch (F) {
#define GSK_PM(V)                                                                                                  \
  case V:                                                                                                          \
    hipLaunchKernelGGL((k_plane_mid<N, V>), dim3(N, F, cnt), dim3(plane_threads<N>()), plane_lds<N>(), st, nk,    \
                       (const float*)nullptr, geo_next(s0), mk, s0->twN, s0->twM, s0->rtn, tab);                   \
    break;
          GSK_PM(1) GSK_PM(2) GSK_PM(3) GSK_PM(4)
#undef GSK_PM
          default: return fail(GSK_ERR_UNSUPPORTED, "batched ensemble launches take 1..4 variants");
        }
        HIPCHK(hipGetLastError());
      }
    } else {
      CHK(y(false));
      switch (F) {
#define GSK_ZM(V)                                                                                                  \
  case V:                                                                                                          \
    hipLaunchKernelGGL((k_zmid_mart<M, V>), dim3(s0->n * s0->n / LZ, 1, cnt), dim3(ZMT), ZM_LDS, st, nk,         \
                       (const float*)nullptr, s0->g, mk, s0->twM, s0->rtn, tab);                                   \
    break;
        GSK_ZM(1) GSK_ZM(2) GSK_ZM(3) GSK_ZM(4)
#undef GSK_ZM
        default: return fail(GSK_ERR_UNSUPPORTED, "batched ensemble launches take 1..4 variants");
      }
      HIPCHK(hipGetLastError());
      CHK(y(true));
    }
    hipLaunchKernelGGL((k_xupdate_mart<N, 1, 0>), dim3(gx, F, cnt), dim3(s0->xthreads), s0->lds_y, st, nk,
                       (float2*)nullptr, geo_next(s0), mk, s0->twN, 0, tab);
    HIPCHK(hipGetLastError());
    if (pl) {
      if constexpr (plane_size<N>()) {
        hipLaunchKernelGGL((k_plane_sq<N>), dim3(N, F, cnt), dim3(plane_threads<N>()), plane_lds<N>(), st, nk,
                           (float*)nullptr, geo_next(s0), s0->twN, s0->twM, s0->rtn, (uint32_t*)nullptr, tab);
      }
    } else {
      CHK(y(false));
      hipLaunchKernelGGL((k_zinv_sq<M>), dim3(rows, F, cnt), dim3(s0->zthreads), lz, st, nk, (float*)nullptr,
                         geo_next(s0), s0->twM, s0->rtn, (uint32_t*)nullptr, tab);
    }
    HIPCHK(hipGetLastError());
    return GSK_OK;
  }

  // ---- 3D transforms of F real fields <-> natural-layout half spectra in an explicit buffer
  static int r2c_buf(gsk_sim* s, const float* in, float2* kb, int F) {
    const int rows = s->n * s->n / s->g.L;
    LAUNCH((k_zfwd<M, Z_COPY>), dim3(rows, F), s->zthreads, zlds(s, 1), s, in, kb, s->g, s->twM, s->rtn);
    LAUNCH((k_ypass<N, true>), dim3(s->g.ntile, F), s->xthreads, s->lds_y, s, kb, s->g, F, s->twN, (const MemberPtrs*)nullptr);
    LAUNCH((k_xpass<N, true>), s->g.ntile, s->xthreads, s->lds_y, s, kb, s->g, F, 1.0f, s->twN);
    return GSK_OK;
  }
  static int c2r_buf(gsk_sim* s, float2* kb, float* out, int F) {
    LAUNCH((k_xpass<N, false>), s->g.ntile, s->xthreads, s->lds_y, s, kb, s->g, F, 1.0f, s->twN);
    LAUNCH((k_ypass<N, false>), dim3(s->g.ntile, F), s->xthreads, s->lds_y, s, kb, s->g, F, s->twN, (const MemberPtrs*)nullptr);
    const int rows = s->n * s->n / s->g.L;
    LAUNCH((k_zinv<M>), dim3(rows, F), s->zthreads, zlds(s, 1), s, kb, out, s->g, s->twM, s->rtn,
           (uint32_t*)nullptr);
    return GSK_OK;
  }

  // One reverse step of the martensite update at eta (step index k.step), eta1 = next state.
  // lam (in: dLoss/deta1, out: dLoss/deta); partial sums into part (nb X, nb*V*V G, nb <mu,F>).
  static int adj_step(gsk_sim* s, const float* eta, const float* eta1, float* lam, float* mu, float* rq, float2* kb,
                      float2* k2, const float* bh, double* part, int nb, const AdjK& k) {
    const int V = s->V;
    const long N3 = (long)s->n * s->n * s->n;
    CHK(r2c_buf(s, lam, kb, V));
    CHK(r2c_buf(s, eta1, k2, V));
    LAUNCH(k_adj_dmul, nb, 256, 0, s, kb, k2, s->g, k, part);
    CHK(c2r_buf(s, kb, mu, V));
    LAUNCH(k_adj_prep, nb, 256, 0, s, eta, mu, rq, V, N3);
    CHK(r2c_buf(s, rq, kb, 2 * V));
    double* pg = part + nb;
    switch (V) {
      case 1: LAUNCH((k_adj_contract<1>), nb, 256, 0, s, kb, bh, s->g, k.inv_n3, pg); break;
      case 2: LAUNCH((k_adj_contract<2>), nb, 256, 0, s, kb, bh, s->g, k.inv_n3, pg); break;
      case 3: LAUNCH((k_adj_contract<3>), nb, 256, 0, s, kb, bh, s->g, k.inv_n3, pg); break;
      case 4: LAUNCH((k_adj_contract<4>), nb, 256, 0, s, kb, bh, s->g, k.inv_n3, pg); break;
      default: return fail(GSK_ERR_UNSUPPORTED, "the adjoint supports up to 4 variants");
    }
    CHK(c2r_buf(s, kb, rq, 2 * V));
    LAUNCH(k_adj_lambda, nb, 256, 0, s, eta, mu, rq, lam, N3, k, part + nb + nb * V * V);
    return GSK_OK;
  }

  // R_p = IFFT[sum_q B_pq FFT(eta_q^2)]/N of the V fields at eta into out (real, V fields)
  static int elastic_R_real(gsk_sim* s, const float* eta, float* out) {
    const int V = s->V, gx = s->g.ntile;
    MartK mk = mart_consts(s);
    CHK(zfwd(s, Z_SQUARE, eta, V));
    CHK(ypass(s, true, V));
    switch (V) {
      case 1: LAUNCH((k_xcontract_r<N, 1>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->Bt, s->g, mk, s->twN, (const MemberPtrs*)nullptr); break;
      case 2: LAUNCH((k_xcontract_r<N, 2>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->Bt, s->g, mk, s->twN, (const MemberPtrs*)nullptr); break;
      case 3: LAUNCH((k_xcontract_r<N, 3>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->Bt, s->g, mk, s->twN, (const MemberPtrs*)nullptr); break;
      case 4: LAUNCH((k_xcontract_r<N, 4>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->Bt, s->g, mk, s->twN, (const MemberPtrs*)nullptr); break;
      default: LAUNCH((k_xcontract<N>), gx, s->xthreads, s->lds_xc, s, s->kbuf, s->Bt, s->g, mk, s->twN);
    }
    CHK(ypass(s, false, V));
    CHK(zinv(s, out, V));
    return GSK_OK;
  }

  static int step_mart_ref(gsk_sim* s) {
    const gsk_config& c = s->cfg;
    CHK(elastic_R_real(s, s->fld, s->rbuf));
    RefMartK k{};
    k.a1 = (float)(0.02 * (c.temperature - c.T0));
    k.a2 = (float)c.B;
    k.a3 = (float)c.C;
    k.sc = (float)std::sqrt(std::fabs((0.02 * (c.temperature - c.T0)) / (c.B - c.C)));
    k.wG = (float)c.w_grad; k.wC = (float)c.w_chem; k.wE = (float)c.w_el; k.wD = (float)c.w_dis;
    k.wN = (float)c.noise; k.dt = (float)c.dt;
    k.g = (float)(1.0 / (2.0 * c.dx) / 3.0);
    k.V = s->V;
    k.chem_form = c.chem_form;
    k.step = (uint32_t)s->step;
    k.stepp = s->dstep;
    k.seed = c.seed;
    const long N3 = (long)s->n * s->n * s->n;
    LAUNCH(k_ref_mart, grid_1d(N3 / 4), 256, 0, s, s->fld, s->rbuf, s->defect, s->Kd, s->n, k);
    LAUNCH(k_copy4, grid_1d((long)s->V * N3 / 4), 256, 0, s, s->fld, s->rbuf, (long)s->V * N3);
    LAUNCH(k_advance, 1, 64, 0, s, s->dstep);
    return GSK_OK;
  }

  static int step_prec_ref(gsk_sim* s) {
    CHK(elastic_R_real(s, s->fld + nvox(s), s->rbuf));
    return step_prec_fused(s);
  }

  // With 2..5 fields the last z pass is fused with the next step's first one (k_zinv_local_r)
  // and the x update advances the noise counter.
  static int step_prec_fused(gsk_sim* s) {
    const int F = s->F;
    PrecK pk = prec_consts(s);
    const bool fuse = prec_carries(s);
    if (!fuse || !s->sq_ready) CHK(zfwd(s, Z_PREC_LOCAL, s->fld, F));
    CHK(ypass(s, true, F));
    if (fuse) pk.adv = s->dstep;
    CHK(xupdate_prec<0>(s, pk));
    CHK(ypass(s, false, F));
    if (!fuse) return zinv(s, s->fld, F, s->dstep);
    pk.adv = nullptr;
    const int rows = s->n * s->n / s->g.L;
    switch (F) {
      case 2: LAUNCH((k_zinv_local_r<M, 2>), rows, s->zthreads, zlds(s, 1), s, s->kbuf, s->fld, s->g, pk, s->twM, s->rtn); break;
      case 3: LAUNCH((k_zinv_local_r<M, 3>), rows, s->zthreads, zlds(s, 1), s, s->kbuf, s->fld, s->g, pk, s->twM, s->rtn); break;
      case 4: LAUNCH((k_zinv_local_r<M, 4>), rows, s->zthreads, zlds(s, 1), s, s->kbuf, s->fld, s->g, pk, s->twM, s->rtn); break;
      default: LAUNCH((k_zinv_local_r<M, 5>), rows, s->zthreads, zlds(s, 1), s, s->kbuf, s->fld, s->g, pk, s->twM, s->rtn);
    }
    s->sq_ready = true;
    return GSK_OK;
  }

  // ---- plain FP64 3D C2C on V fields of buf
  static int c2c(gsk_sim* s, double2* buf, bool fwd) {
    const long n = s->n;
    const long N3 = n * n * n;
    for (int f = 0; f < s->V; ++f) {
      double2* b = buf + f * N3;
      const int gz = (int)(n * n / s->pL);
      if (fwd) LAUNCH((p_zrows<N, true>), gz, s->pzthreads, s->lds_pz, s, b, s->pL, s->twN64);
      else LAUNCH((p_zrows<N, false>), gz, s->pzthreads, s->lds_pz, s, b, s->pL, s->twN64);
      const int gs = (int)(n * (n / s->pW));
      // y: outer x (stride n^2), positions y (stride n)
      if (fwd) LAUNCH((p_strided<N, true>), gs, s->pxthreads, s->lds_px, s, b, s->pW, n * n, n, s->twN64);
      else LAUNCH((p_strided<N, false>), gs, s->pxthreads, s->lds_px, s, b, s->pW, n * n, n, s->twN64);
      // x: outer y (stride n), positions x (stride n^2)
      if (fwd) LAUNCH((p_strided<N, true>), gs, s->pxthreads, s->lds_px, s, b, s->pW, n, n * n, s->twN64);
      else LAUNCH((p_strided<N, false>), gs, s->pxthreads, s->lds_px, s, b, s->pW, n, n * n, s->twN64);
    }
    return GSK_OK;
  }

  static int step_mart_plain(gsk_sim* s) {
    const long N3 = (long)nvox(s), VN = s->V * N3;
    PlainK k = plain_consts(s);
    LAUNCH(p_square, grid_1d(VN), 256, 0, s, s->eta64, s->S64, VN);
    CHK(c2c(s, s->S64, true));
    LAUNCH(p_contract, grid_1d(N3), 256, 0, s, s->S64, s->R64, s->B64, N3, k);
    CHK(c2c(s, s->R64, false));
    LAUNCH(p_force, grid_1d(N3 / 4), 256, 0, s, s->eta64, s->R64, s->F64, N3, k);
    LAUNCH(p_load, grid_1d(VN), 256, 0, s, s->eta64, s->E64, VN);
    CHK(c2c(s, s->F64, true));
    CHK(c2c(s, s->E64, true));
    LAUNCH(p_update, grid_1d(VN), 256, 0, s, s->E64, s->F64, s->n, k);
    CHK(c2c(s, s->E64, false));
    LAUNCH(p_extract, grid_1d(VN), 256, 0, s, s->E64, s->eta64, VN);
    return GSK_OK;
  }

  static int energy_mart_el_fused(gsk_sim* s, double* el) {
    // S_hat = FFT(eta^2) through z and y; the energy kernel applies x and contracts.
    if (plane_on(s)) {
      if (!s->sq_ready) CHK(square_fwd(s));  // y,z-transformed already
    } else {
      if (!s->sq_ready) CHK(zfwd(s, Z_SQUARE, s->fld, s->F));
      CHK(ypass(s, true, s->F));
    }
    s->sq_ready = false;  // the buffer is no longer what a step starts from
    const int gx = s->g.ntile;
    if (gx > s->rblocks) return fail(GSK_ERR, "reduction scratch too small");
    switch (s->F) {
      case 1: LAUNCH((r_xenergy_mart_r<N, 1>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->Bt, s->g, s->twN, s->part); break;
      case 2: LAUNCH((r_xenergy_mart_r<N, 2>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->Bt, s->g, s->twN, s->part); break;
      case 3: LAUNCH((r_xenergy_mart_r<N, 3>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->Bt, s->g, s->twN, s->part); break;
      case 4: LAUNCH((r_xenergy_mart_r<N, 4>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->Bt, s->g, s->twN, s->part); break;
      default: {
        const size_t lds = (size_t)(N + (size_t)s->F * N * s->g.W) * sizeof(float2);
        if (lds > 160 * 1024) return fail(GSK_ERR_UNSUPPORTED, "elastic energy: too many variants for this grid");
        LAUNCH((r_xenergy_mart<N>), gx, s->xthreads, lds, s, s->kbuf, s->Bt, s->g, s->F, s->twN, s->part);
      }
    }
    LAUNCH(r_finish, 1, 1024, 0, s, s->part, gx, 1, s->rout);
    HIPCHK(hipMemcpyAsync(el, s->rout, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return GSK_OK;
  }

  static int energy_mart_el_plain(gsk_sim* s, double* gr, double* el) {
    const long N3 = (long)nvox(s), VN = s->V * N3;
    PlainK k = plain_consts(s);
    LAUNCH(p_square, grid_1d(VN), 256, 0, s, s->eta64, s->S64, VN);
    CHK(c2c(s, s->S64, true));
    LAUNCH(p_load, grid_1d(VN), 256, 0, s, s->eta64, s->E64, VN);
    CHK(c2c(s, s->E64, true));
    const int nb = 256;
    LAUNCH(r_plain_kspace, nb, 256, 0, s, s->E64, s->S64, s->B64, s->n, s->V, k.kfac, s->part);
    LAUNCH(r_finish, 1, 1024, 0, s, s->part, nb, 2, s->rout);
    double h[2];
    HIPCHK(hipMemcpyAsync(h, s->rout, 2 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    *gr = h[0];
    *el = h[1];
    return GSK_OK;
  }
};

template <template <int> class Fn, typename... Args>
int dispatch_n(int n, Args&&... args) {
  switch (n) {
    case 16: return Fn<16>::run(args...);
    case 32: return Fn<32>::run(args...);
    case 64: return Fn<64>::run(args...);
    case 96: return Fn<96>::run(args...);
    case 128: return Fn<128>::run(args...);
    case 192: return Fn<192>::run(args...);
    case 256: return Fn<256>::run(args...);
    case 512: return Fn<512>::run(args...);
  }
  return fail(GSK_ERR_UNSUPPORTED, "unsupported grid size " + std::to_string(n));
}

template <int N> struct DAttrs { static int run() { return Impl<N>::set_attrs(); } };
template <int N> struct DInit { static int run(gsk_sim* s) { return Impl<N>::init_state(s); } };
template <int N> struct DSquare { static int run(gsk_sim* s) { return Impl<N>::square_fwd(s); } };
template <int N> struct DBatch {
  static int run(gsk_sim* s0, int cnt, const MemberPtrs* tab, hipStream_t st) {
    return Impl<N>::step_mart_batched(s0, cnt, tab, st);
  }
};
template <int N> struct DStep {
  static int run(gsk_sim* s) {
    if (s->cfg.path == GSK_PATH_PLAIN64) return Impl<N>::step_mart_plain(s);
    if (s->cfg.rule == GSK_RULE_REFERENCE)
      return s->cfg.physics == GSK_PHYS_MART ? Impl<N>::step_mart_ref(s) : Impl<N>::step_prec_ref(s);
    if (s->cfg.physics == GSK_PHYS_MART) return Impl<N>::step_mart_fused(s);
    return Impl<N>::step_prec_fused(s);
  }
};
template <int N> struct DElF { static int run(gsk_sim* s, double* el) { return Impl<N>::energy_mart_el_fused(s, el); } };
template <int N> struct DElP {
  static int run(gsk_sim* s, double* gr, double* el) { return Impl<N>::energy_mart_el_plain(s, gr, el); }
};
template <int N> struct DZfwd { static int run(gsk_sim* s, const float* in, int F) { return Impl<N>::zfwd(s, Z_COPY, in, F); } };
template <int N> struct DY { static int run(gsk_sim* s, bool fwd, int F) { return Impl<N>::ypass(s, fwd, F); } };
template <int N> struct DX {
  static int run(gsk_sim* s, bool fwd, int F, float scale) {
    const int gx = s->g.ntile;
    if (fwd) LAUNCH((k_xpass<N, true>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->g, F, scale, s->twN);
    else LAUNCH((k_xpass<N, false>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->g, F, scale, s->twN);
    return GSK_OK;
  }
};
template <int N> struct DAdj {
  static int run(gsk_sim* s, const float* eta, const float* eta1, float* lam, float* mu, float* rq, float2* kb,
                 float2* k2, const float* bh, double* part, int nb, const AdjK& k) {
    return Impl<N>::adj_step(s, eta, eta1, lam, mu, rq, kb, k2, bh, part, nb, k);
  }
};
template <int N> struct DR2C {
  static int run(gsk_sim* s, const float* in, float2* kb, int F) { return Impl<N>::r2c_buf(s, in, kb, F); }
};
template <int N> struct DC2R {
  static int run(gsk_sim* s, float2* kb, float* out, int F) { return Impl<N>::c2r_buf(s, kb, out, F); }
};
template <int N> struct DZinv { static int run(gsk_sim* s, float* out, int F) { return Impl<N>::zinv(s, out, F); } };

int ensure_device(int dev) {
  int cnt = 0;
  HIPCHK(hipGetDeviceCount(&cnt));
  if (cnt <= 0) return fail(GSK_ERR_HIP, "no HIP device available");
  if (dev < 0 || dev >= cnt) return fail(GSK_ERR_ARG, "device ordinal out of range");
  HIPCHK(hipSetDevice(dev));
  return GSK_OK;
}

// Choose tile shapes for a given field count F.
void choose_geometry(gsk_sim* s, int Fz, int Fx_resident) {
  const int n = s->n, M = n / 2;
  Geo& g = s->g;
  g.n = n; g.M = M;
  g.nvox = (long long)n * n * n;
  g.nq = (long long)n * n * M;
  g.kfs = ((long long)n * n * (M + 1) + 15) / 16 * 16;
  // z tiles: the compile-time shape of the z kernels (zrows rows, zthreads threads)
  const int L = zrows(M);
  g.L = L;
  s->zthreads = zthreads_host(M);
  // x/y tiles: width W <= 16 (GSK_W overrides) dividing n/2, threads W*n/rmin <= 1024,
  // resident tiles <= 120 KB
  const int rx = rmin_of(n);
  int W = 16;
  // (the strided kernels are compiled for at most 16 lines per tile, xy_threads<N>())
  if (const char* e = std::getenv("GSK_W")) W = std::max(1, std::min(16, std::atoi(e)));
  while (W > 1 && (M % W || W * n / rx > 1024 || (double)(n + (double)Fx_resident * n * W) * 8.0 > 120 * 1024))
    W /= 2;
  g.W = W;
  g.nchm = M / W;
  g.ntile = n * g.nchm + n / W;
  s->xthreads = round64(std::max(W * n / rx, 64));
  s->lds_y = (size_t)(n + n * W) * sizeof(float2);
  s->lds_xc = (size_t)(n + (size_t)Fx_resident * n * W) * sizeof(float2);
  s->lds_z = (2 * M + (size_t)Fz * L * zpitch(M)) * sizeof(float2);
}

void choose_plain(gsk_sim* s) {
  const int n = s->n, r = rmin_of(n);
  int L = 1;
  while (L * 2 <= n * n && (L * 2) * n / r <= 256 && (double)(n + (L * 2) * (n + 1)) * 16.0 <= 96 * 1024) L *= 2;
  s->pL = L;
  s->pzthreads = round64(std::max(L * n / r, 64));
  int W = 16;
  while (W > 1 && (W * n / r > 1024 || (double)(n + n * W) * 16.0 > 96 * 1024 || n % W)) W /= 2;
  s->pW = W;
  s->pxthreads = round64(std::max(W * n / r, 64));
  s->lds_pz = (size_t)(n + L * (n + 1)) * sizeof(double2);
  s->lds_px = (size_t)(n + n * W) * sizeof(double2);
}

void build_slots(gsk_sim* s) {
  s->slots.clear();
  const double hs = (double)hs_count(s) * 8.0;  // one half-spectrum complex field
  const double rv = (double)nvox(s) * 4.0;      // one real fp32 field
  const int F = s->F;
  if (s->cfg.path == GSK_PATH_PLAIN64) {
    const double cf = (double)nvox(s) * 16.0, r8 = (double)nvox(s) * 8.0;
    const int V = s->V;
    auto add_c2c = [&](const char* a, const char* b, const char* c) {
      for (int f = 0; f < V; ++f) {
        s->slots.push_back({a, 2 * cf});
        s->slots.push_back({b, 2 * cf});
        s->slots.push_back({c, 2 * cf});
      }
    };
    s->slots.push_back({"p_square", V * (r8 + cf)});
    add_c2c("p_z_fwdS", "p_y_fwdS", "p_x_fwdS");
    s->slots.push_back({"p_contract", V * 2 * cf + V * V * r8});
    add_c2c("p_z_invR", "p_y_invR", "p_x_invR");
    s->slots.push_back({"p_force", V * (r8 + cf + cf)});
    s->slots.push_back({"p_load", V * (r8 + cf)});
    add_c2c("p_z_fwdF", "p_y_fwdF", "p_x_fwdF");
    add_c2c("p_z_fwdE", "p_y_fwdE", "p_x_fwdE");
    s->slots.push_back({"p_update", V * 3 * cf});
    add_c2c("p_z_invE", "p_y_invE", "p_x_invE");
    s->slots.push_back({"p_extract", V * (cf + r8)});
    return;
  }
  if (s->cfg.physics == GSK_PHYS_MART) {
    const int npq = F * (F + 1) / 2;
    const bool pl = plane_on(s);
    if (!pl) s->slots.push_back({"y_fwd_S", F * 2 * hs});
    s->slots.push_back({"x_contract", F * 2 * hs + npq * hs / 2});
    if (pl) {  // R and F spectra, eta of every variant per variant
      s->slots.push_back({"plane_mid", F * (2 * hs + F * rv)});
    } else {
      s->slots.push_back({"y_inv_R", F * 2 * hs});
      s->slots.push_back({"zmid_force", F * (2 * hs + rv)});
      s->slots.push_back({"y_fwd_F", F * 2 * hs});
    }
    static const char* const xname[] = {"x_update_g0", "x_update_g1", "x_update_g2", "x_update_g3",
                                        "x_update_g4", "x_update_g5", "x_update_g6", "x_update_g7"};
    for (int f0 = 0, gi = 0; f0 < F; ++gi) {  // one launch per field group, as Impl::xupdate_mart
      const int c = xupd_group(F - f0);
      s->slots.push_back({xname[std::min(gi, 7)], c * 4 * hs});
      f0 += c;
    }
    if (pl) {
      s->slots.push_back({"plane_sq", F * (2 * hs + rv)});  // eta and S spectra, eta written
    } else {
      s->slots.push_back({"y_inv_eta", F * 2 * hs});
      s->slots.push_back({"zinv_eta_sq", F * (2 * hs + rv)});  // C2R, write eta, R2C of eta^2
    }
  } else {
    const bool fuse = prec_carries(s);
    if (!fuse) s->slots.push_back({"zfwd_local", F * (rv + hs)});
    s->slots.push_back({"y_fwd", F * 2 * hs});
    if (GSK_XUPD_G == 0) {  // one launch, as Impl::xupdate_prec
      s->slots.push_back({"x_update", F * 4 * hs + hs / 2});
    } else {
      for (int f0 = 0; f0 < F; f0 += 2) {
        const int c = std::min(2, F - f0);
        s->slots.push_back({f0 == 0 ? "x_update_g0" : (f0 == 2 ? "x_update_g1" : (f0 == 4 ? "x_update_g2" : "x_update_g3")),
                            c * 4 * hs + (f0 == 0 ? hs / 2 : 0.0)});
      }
    }
    s->slots.push_back({"y_inv", F * 2 * hs});
    if (fuse) s->slots.push_back({"zinv_local", F * (2 * hs + rv)});  // C2R, write fields, local forces, R2C
    else s->slots.push_back({"zinv", F * (hs + rv)});
  }
}

int sync(gsk_sim* s) {
  HIPCHK(hipStreamSynchronize(s->stream));
  return GSK_OK;
}

int prepare(gsk_sim* s) {
  if (!s->haveB) return fail(GSK_ERR_STATE, "B tensor not built (call gsk_build_btensor first)");
  const bool stateless = s->cfg.rule == GSK_RULE_REFERENCE && s->cfg.physics == GSK_PHYS_MART;
  if (s->dirty && s->cfg.path == GSK_PATH_FUSED32 && !stateless) {
    s->sq_ready = false;  // the initialisation uses the k-space buffer
    CHK(dispatch_n<DInit>(s->n, s));
  }
  s->dirty = false;
  // the spectral steps start from the z-transformed eta^2 (martensite) or local forces
  // (precipitate) in k-space: computed here, so that every step of a timed or profiled run
  // has the same launches
  if (carries(s) && !s->sq_ready) CHK(dispatch_n<DSquare>(s->n, s));
  return GSK_OK;
}

template <typename T>
void free_dev(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

void adj_free(AdjWS* w);

}  // namespace

// ======================================================================== C ABI
extern "C" {

const char* gsk_last_error(void) { return g_err.c_str(); }

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
  if (cfg->physics != GSK_PHYS_MART && cfg->physics != GSK_PHYS_PREC) return fail(GSK_ERR_ARG, "unknown physics");
  if (cfg->path != GSK_PATH_FUSED32 && cfg->path != GSK_PATH_PLAIN64) return fail(GSK_ERR_ARG, "unknown path");
  if (cfg->path == GSK_PATH_PLAIN64 && cfg->physics != GSK_PHYS_MART)
    return fail(GSK_ERR_UNSUPPORTED, "the FP64 plain path implements the martensite rule only");
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
  if (cfg->path == GSK_PATH_FUSED32) {
    if ((rc = alloc(s->fld, (size_t)s->F * N3)) != GSK_OK) return bail(rc);
    if ((rc = alloc(s->kbuf, (size_t)s->F * s->g.kfs)) != GSK_OK) return bail(rc);
    if ((rc = alloc(s->state, ntile * s->F * T)) != GSK_OK) return bail(rc);
    const bool vv = cfg->physics == GSK_PHYS_MART || cfg->rule == GSK_RULE_REFERENCE;
    const int npq = vv ? s->V * (s->V + 1) / 2 : 1;
    if ((rc = alloc(s->Bt, ntile * npq * T)) != GSK_OK) return bail(rc);
    if (cfg->rule == GSK_RULE_REFERENCE && (rc = alloc(s->rbuf, (size_t)s->V * N3)) != GSK_OK) return bail(rc);
  } else {
    const size_t VN = (size_t)s->V * N3;
    if ((rc = alloc(s->eta64, VN)) != GSK_OK) return bail(rc);
    if ((rc = alloc(s->S64, VN)) != GSK_OK) return bail(rc);
    if ((rc = alloc(s->R64, VN)) != GSK_OK) return bail(rc);
    if ((rc = alloc(s->E64, VN)) != GSK_OK) return bail(rc);
    if ((rc = alloc(s->F64, VN)) != GSK_OK) return bail(rc);
    if ((rc = alloc(s->B64, (size_t)s->V * s->V * N3)) != GSK_OK) return bail(rc);
  }
  s->rblocks = std::max<int>(1024, (int)ntile);
  if ((rc = alloc(s->part, (size_t)s->rblocks * 16)) != GSK_OK) return bail(rc);
  if ((rc = alloc(s->rout, 16)) != GSK_OK) return bail(rc);
  if (cfg->path == GSK_PATH_FUSED32 && (rc = alloc(s->dstep, 1)) != GSK_OK) return bail(rc);
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
  free_dev(s->twN); free_dev(s->twM); free_dev(s->rtn); free_dev(s->twN64);
  if (s->sharedB) s->Bt = nullptr;
  free_dev(s->fld); free_dev(s->kbuf); free_dev(s->state); free_dev(s->Bt); free_dev(s->dstep);
  free_dev(s->eta64); free_dev(s->S64); free_dev(s->R64); free_dev(s->E64); free_dev(s->F64); free_dev(s->B64);
  free_dev(s->part); free_dev(s->rout);
  free_dev(s->rbuf); free_dev(s->defect); free_dev(s->Kd);
  for (auto& e : s->ev) (void)hipEventDestroy(e);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return GSK_OK;
}

static int fill_bargs(BArgs& a, const double* C81, const double* e1, int V1, const double* e2, int V2, double kfac) {
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
  if (s->cfg.path == GSK_PATH_FUSED32) {
    const long total = (long)(long)s->g.ntile * s->n * s->g.W;
    hipLaunchKernelGGL(k_btensor_tiles, dim3(grid_1d(total)), dim3(256), 0, s->stream, s->Bt, s->g, a);
  } else {
    hipLaunchKernelGGL(k_btensor_full, dim3(grid_1d((long)nvox(s))), dim3(256), 0, s->stream, s->B64, s->n, a);
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
  if (s->cfg.path == GSK_PATH_PLAIN64) {
    std::vector<double> h(count);
    for (size_t i = 0; i < count; ++i)
      h[i] = dtype == GSK_F32 ? (double)((const float*)host)[i] : ((const double*)host)[i];
    CHK(sync(s));  // earlier work on the handle's stream may still read the field
    HIPCHK(hipMemcpy(s->eta64, h.data(), count * sizeof(double), hipMemcpyHostToDevice));
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
  if (s->cfg.path == GSK_PATH_PLAIN64) {
    std::vector<double> h(count);
    HIPCHK(hipMemcpy(h.data(), s->eta64, count * sizeof(double), hipMemcpyDeviceToHost));
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
    hipLaunchKernelGGL(k_uniform, dim3(grid_1d((long)N3 / 4)), dim3(256), 0, s->stream, tmp + f * N3, (long)N3, seed,
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
    CHK(dispatch_n<DStep>(s->n, s));
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
  auto real_pass = [&]() -> int {
    if (s->cfg.path == GSK_PATH_PLAIN64) {
      LAUNCH(r_mart_real<double>, nb, 256, 0, s, s->eta64, N3, s->V, c.A, c.B, c.C, s->part);
    } else if (c.physics == GSK_PHYS_MART) {
      LAUNCH(r_mart_real<float>, nb, 256, 0, s, s->fld, N3, s->V, c.A, c.B, c.C, s->part);
    } else {
      PrecE pe{};
      for (int i = 0; i < 8; ++i) pe.A[i] = c.a[i];
      pe.c1 = c.c1; pe.c2 = c.c2; pe.cbar = s->cbar; pe.V = s->V;
      LAUNCH(r_prec_real, nb, 256, 0, s, s->fld, N3, pe, s->part);
    }
    LAUNCH(r_finish, 1, 1024, 0, s, s->part, nb, 2 + s->V, s->rout);
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
  if (what == GSK_RED_MASS) {
    if (c.physics != GSK_PHYS_PREC) return fail(GSK_ERR_UNSUPPORTED, "mass is a precipitate reduction");
    CHK(prepare(s));
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
    if (s->cfg.path == GSK_PATH_PLAIN64) {
      double g2, e2;
      CHK(dispatch_n<DElP>(s->n, s, &g2, &e2));
      gr = 0.5 * c.kappa * g2 / (double)N3;
      el = 0.5 * c.w_el * e2 / (double)N3;
    } else {
      const int ntile = s->g.ntile;
      const double kfac = c.kconv == GSK_K_REFERENCE ? 2.0 * 3.14 * c.dx / s->n : 2.0 * kPi / (s->n * c.dx);
      LAUNCH(r_state_grad, nb, 256, 0, s, s->state, s->g, s->F, (float)kfac, s->part);
      LAUNCH(r_finish, 1, 1024, 0, s, s->part, nb, s->F, s->rout);
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
        LAUNCH(r_state_el_prec, nb, 256, 0, s, s->state, s->Bt, s->g, s->F, s->part);
        LAUNCH(r_finish, 1, 1024, 0, s, s->part, nb, 1, s->rout);
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
    CHK(dispatch_n<DStep>(s->n, s));
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
  hipLaunchKernelGGL(k_noise, dim3(grid_1d(nv / 4)), dim3(256), 0, 0, d, (long)nv, seed, step, field);
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
  hipLaunchKernelGGL(k_btensor_full, dim3(grid_1d((long)N3)), dim3(256), 0, 0, d, n, a);
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

// ======================================================================== ensembles
// Many independent instances (any mix of grid sizes and parameters) stepped together.
// Members keep their own buffers; instances with the same grid, variant count and elastic
// data share one B tensor. Three schedules:
//   GSK_ENS_SEQUENTIAL  each member runs its steps alone, one after another;
//   GSK_ENS_STREAMS     members are spread over `branches` streams (greedy balance of
//                       voxels, largest first) and launched eagerly;
//   GSK_ENS_GRAPH       the same branches captured once as a hipGraph holding
//                       `steps_per_graph` steps of every member, replayed as a whole.
// The noise counter of every member lives on the device and is advanced by the last
// kernel of each step, so one graph serves every step.
struct gsk_ensemble {
  std::vector<gsk_sim*> m;
  std::vector<float*> sharedB;
  int dev = 0;
  int mode = GSK_ENS_GRAPH, branches = 0, spg = 10;
  hipStream_t master = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipStream_t> bs;
  std::vector<hipEvent_t> join;
  hipGraphExec_t exec = nullptr;
  int g_mode = -1, g_branches = -1, g_spg = -1;
  std::vector<unsigned> g_gen;
  // batched schedules: members of one grid size and variant count share every launch
  struct Cls {
    std::vector<int> idx;          // members of the class
    MemberPtrs* dtab = nullptr;    // their device table (buffers and constants)
  };
  std::vector<Cls> cls;
  std::vector<int> loose;          // members that run one by one (other rules, physics or V > 4)
  std::vector<unsigned> cls_gen;   // member generations the tables were built from
};

namespace {

int ens_branches(const gsk_ensemble* e) {
  const int cnt = (int)e->m.size();
  const int b = e->branches > 0 ? e->branches : std::min(cnt, 8);
  return std::max(1, std::min(b, cnt));
}

// members per branch: greedy balance of voxel-fields, largest members first
std::vector<std::vector<int>> ens_assign(const gsk_ensemble* e, int B) {
  std::vector<int> idx(e->m.size());
  for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
  auto work = [&](int i) { return (double)nvox(e->m[i]) * e->m[i]->F; };
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return work(a) > work(b); });
  std::vector<double> load(B, 0.0);
  std::vector<std::vector<int>> out(B);
  for (int i : idx) {
    const int b = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    out[b].push_back(i);
    load[b] += work(i);
  }
  return out;
}

int ens_streams(gsk_ensemble* e, int B) {
  while ((int)e->bs.size() < B) {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    e->bs.push_back(st);
    HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    e->join.push_back(ev);
  }
  return GSK_OK;
}

int sync_all(gsk_ensemble* e) {
  if (e->master) HIPCHK(hipStreamSynchronize(e->master));
  for (hipStream_t st : e->bs) HIPCHK(hipStreamSynchronize(st));
  for (gsk_sim* s : e->m) CHK(sync(s));
  return GSK_OK;
}

int ens_prepare(gsk_ensemble* e) {
  CHK(ensure_device(e->dev));
  for (gsk_sim* s : e->m) {
    if (s->prof) return fail(GSK_ERR_STATE, "an ensemble member is being profiled");
    if (s->cfg.path != GSK_PATH_FUSED32) return fail(GSK_ERR_UNSUPPORTED, "ensembles run the FP32 fused path");
    CHK(prepare(s));
  }
  for (gsk_sim* s : e->m) CHK(sync(s));
  return GSK_OK;
}

// Enqueue `steps` steps of every member on the branch streams, forked from and joined
// back into the master stream (eagerly, or into a graph when the master is capturing).
int ens_launch(gsk_ensemble* e, const std::vector<std::vector<int>>& asg, int steps) {
  const int B = (int)asg.size();
  HIPCHK(hipEventRecord(e->fork, e->master));
  for (int b = 0; b < B; ++b) HIPCHK(hipStreamWaitEvent(e->bs[b], e->fork, 0));
  for (int b = 0; b < B; ++b)
    for (int i : asg[b]) {
      gsk_sim* s = e->m[i];
      hipStream_t own = s->stream;
      s->stream = e->bs[b];
      int rc = GSK_OK;
      for (int k = 0; k < steps && rc == GSK_OK; ++k) rc = dispatch_n<DStep>(s->n, s);
      s->stream = own;
      CHK(rc);
    }
  for (int b = 0; b < B; ++b) {
    HIPCHK(hipEventRecord(e->join[b], e->bs[b]));
    HIPCHK(hipStreamWaitEvent(e->master, e->join[b], 0));
  }
  return GSK_OK;
}

void ens_drop_graph(gsk_ensemble* e) {
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  e->exec = nullptr;
  e->g_mode = e->g_branches = e->g_spg = -1;
  e->g_gen.clear();
}

// Largest grid edge batched into shared launches. Small grids are launch-bound, so one launch
// for a whole size class pays; larger members run better interleaved on the branch streams
// (96 members of 64/128/192: batching 64 only 16.2 ms per ensemble step, also 128 16.7 ms, every
// size 17.3 ms, streams alone 16.6 ms).
#ifndef GSK_ENS_BATCH_MAXN
#define GSK_ENS_BATCH_MAXN 96
#endif

bool batchable(const gsk_sim* s) {
  return s->cfg.path == GSK_PATH_FUSED32 && s->cfg.physics == GSK_PHYS_MART && s->cfg.rule == GSK_RULE_SPECTRAL &&
         s->V >= 1 && s->V <= 4;
}

void ens_drop_classes(gsk_ensemble* e) {
  for (auto& c : e->cls)
    if (c.dtab) (void)hipFree(c.dtab);
  e->cls.clear();
  e->loose.clear();
  e->cls_gen.clear();
}

// Size classes of the batched schedules and their device tables, rebuilt when a member's
// constants change (the tables hold each member's buffers and physics constants).
int ens_classes(gsk_ensemble* e) {
  std::vector<unsigned> gen;
  for (gsk_sim* s : e->m) gen.push_back(s->gen);
  if (!e->cls_gen.empty() && gen == e->cls_gen) return GSK_OK;
  ens_drop_graph(e);  // a captured graph refers to the old tables
  CHK(sync_all(e));
  ens_drop_classes(e);
  std::vector<std::pair<int, int>> keys;  // (n, V) of each class
  for (int i = 0; i < (int)e->m.size(); ++i) {
    const gsk_sim* s = e->m[i];
    if (!batchable(s) || s->n > GSK_ENS_BATCH_MAXN) {
      e->loose.push_back(i);
      continue;
    }
    size_t c = 0;
    while (c < keys.size() && keys[c] != std::make_pair(s->n, s->V)) ++c;
    if (c == keys.size()) {
      keys.push_back({s->n, s->V});
      e->cls.emplace_back();
    }
    e->cls[c].idx.push_back(i);
  }
  for (auto& c : e->cls) {
    std::vector<MemberPtrs> h;
    for (int i : c.idx) {
      gsk_sim* s = e->m[i];
      h.push_back(MemberPtrs{s->kbuf, s->state, s->fld, s->Bt, s->dstep, mart_consts(s)});
    }
    HIPCHK(hipMalloc((void**)&c.dtab, h.size() * sizeof(MemberPtrs)));
    HIPCHK(hipMemcpy(c.dtab, h.data(), h.size() * sizeof(MemberPtrs), hipMemcpyHostToDevice));
  }
  e->cls_gen = gen;
  return GSK_OK;
}

// Enqueue `steps` batched steps: one branch stream per size class (every launch covers the
// class's members), the other members spread over further branches as in ens_launch.
int ens_launch_batched(gsk_ensemble* e, int steps) {
  const int C = (int)e->cls.size();
  const int Lb = e->loose.empty() ? 0 : std::min((int)e->loose.size(), ens_branches(e));
  CHK(ens_streams(e, C + Lb));
  HIPCHK(hipEventRecord(e->fork, e->master));
  for (int b = 0; b < C + Lb; ++b) HIPCHK(hipStreamWaitEvent(e->bs[b], e->fork, 0));
  for (int c = 0; c < C; ++c) {
    gsk_sim* s0 = e->m[e->cls[c].idx[0]];
    const int cnt = (int)e->cls[c].idx.size();
    for (int k = 0; k < steps; ++k) CHK(dispatch_n<DBatch>(s0->n, s0, cnt, (const MemberPtrs*)e->cls[c].dtab, e->bs[c]));
  }
  for (int j = 0; j < (int)e->loose.size(); ++j) {
    gsk_sim* s = e->m[e->loose[j]];
    hipStream_t own = s->stream;
    s->stream = e->bs[C + j % Lb];
    int rc = GSK_OK;
    for (int k = 0; k < steps && rc == GSK_OK; ++k) rc = dispatch_n<DStep>(s->n, s);
    s->stream = own;
    CHK(rc);
  }
  for (int b = 0; b < C + Lb; ++b) {
    HIPCHK(hipEventRecord(e->join[b], e->bs[b]));
    HIPCHK(hipStreamWaitEvent(e->master, e->join[b], 0));
  }
  return GSK_OK;
}

bool batched_mode(int mode) { return mode == GSK_ENS_BATCHED || mode == GSK_ENS_BATCHED_GRAPH; }

int ens_graph(gsk_ensemble* e, const std::vector<std::vector<int>>& asg, int spg) {
  std::vector<unsigned> gen;
  for (gsk_sim* s : e->m) gen.push_back(s->gen);
  if (e->exec && e->g_mode == e->mode && e->g_branches == (int)asg.size() && e->g_spg == spg && gen == e->g_gen)
    return GSK_OK;
  ens_drop_graph(e);
  HIPCHK(hipStreamBeginCapture(e->master, hipStreamCaptureModeThreadLocal));
  const int rc = batched_mode(e->mode) ? ens_launch_batched(e, spg) : ens_launch(e, asg, spg);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(e->master, &graph);
  if (rc != GSK_OK || ec != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc != GSK_OK ? rc : fail(GSK_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ec));
  }
  const hipError_t ei = hipGraphInstantiate(&e->exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) {
    e->exec = nullptr;
    return fail(GSK_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ei));
  }
  e->g_mode = e->mode;
  e->g_branches = (int)asg.size();
  e->g_spg = spg;
  e->g_gen = gen;
  return GSK_OK;
}

int ens_run(gsk_ensemble* e, int nsteps) {
  CHK(ens_prepare(e));
  if (nsteps == 0) return GSK_OK;
  if (e->mode == GSK_ENS_SEQUENTIAL) {
    for (gsk_sim* s : e->m) {
      for (int k = 0; k < nsteps; ++k) CHK(dispatch_n<DStep>(s->n, s));
      CHK(sync(s));
    }
  } else {
    const bool bat = batched_mode(e->mode);
    const int B = ens_branches(e);
    CHK(ens_streams(e, B));
    if (bat) CHK(ens_classes(e));
    const auto asg = ens_assign(e, B);
    int left = nsteps;
    if ((e->mode == GSK_ENS_GRAPH || e->mode == GSK_ENS_BATCHED_GRAPH) && nsteps >= e->spg) {
      CHK(ens_graph(e, asg, e->spg));
      for (; left >= e->spg; left -= e->spg) HIPCHK(hipGraphLaunch(e->exec, e->master));
    }
    if (left > 0) CHK(bat ? ens_launch_batched(e, left) : ens_launch(e, asg, left));
    HIPCHK(hipStreamSynchronize(e->master));
  }
  for (gsk_sim* s : e->m) s->step += nsteps;
  return GSK_OK;
}

// device copies of the evolving state of every member (real fields, k-space state, step)
struct EnsSnapshot {
  std::vector<float*> fld;
  std::vector<float2*> st;
  std::vector<long long> step;
  ~EnsSnapshot() {
    for (float* p : fld) (void)hipFree(p);
    for (float2* p : st) (void)hipFree(p);
  }
};
size_t state_count(const gsk_sim* s) { return (size_t)s->g.ntile * s->F * s->n * s->g.W; }

int ens_save(gsk_ensemble* e, EnsSnapshot& snap) {
  for (gsk_sim* s : e->m) {
    float* f = nullptr;
    float2* t = nullptr;
    HIPCHK(hipMalloc((void**)&f, (size_t)s->F * nvox(s) * sizeof(float)));
    snap.fld.push_back(f);
    HIPCHK(hipMalloc((void**)&t, state_count(s) * sizeof(float2)));
    snap.st.push_back(t);
    snap.step.push_back(s->step);
    HIPCHK(hipMemcpyAsync(f, s->fld, (size_t)s->F * nvox(s) * sizeof(float), hipMemcpyDeviceToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(t, s->state, state_count(s) * sizeof(float2), hipMemcpyDeviceToDevice, s->stream));
    CHK(sync(s));
  }
  return GSK_OK;
}

int ens_restore(gsk_ensemble* e, const EnsSnapshot& snap) {
  for (size_t i = 0; i < e->m.size(); ++i) {
    gsk_sim* s = e->m[i];
    CHK(sync(s));
    HIPCHK(hipMemcpyAsync(s->fld, snap.fld[i], (size_t)s->F * nvox(s) * sizeof(float), hipMemcpyDeviceToDevice,
                          s->stream));
    HIPCHK(hipMemcpyAsync(s->state, snap.st[i], state_count(s) * sizeof(float2), hipMemcpyDeviceToDevice, s->stream));
    const uint32_t v = (uint32_t)snap.step[i];
    HIPCHK(hipMemcpyAsync(s->dstep, &v, sizeof(v), hipMemcpyHostToDevice, s->stream));
    CHK(sync(s));
    s->step = snap.step[i];
    s->sq_ready = false;  // restored fields: the next step transforms eta^2 afresh
  }
  return GSK_OK;
}

}  // namespace

extern "C" {

int gsk_ensemble_create(const gsk_config* cfgs, int count, gsk_ensemble** out) {
  if (!cfgs || !out || count < 1) return fail(GSK_ERR_ARG, "bad argument");
  *out = nullptr;
  for (int i = 1; i < count; ++i)
    if (cfgs[i].device != cfgs[0].device) return fail(GSK_ERR_ARG, "ensemble members must share one device");
  CHK(ensure_device(cfgs[0].device));
  gsk_ensemble* e = new gsk_ensemble();
  e->dev = cfgs[0].device;
  auto bail = [&](int rc) { gsk_ensemble_destroy(e); return rc; };
  if (hipStreamCreateWithFlags(&e->master, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) != hipSuccess)
    return bail(fail(GSK_ERR_HIP, "stream/event creation failed"));
  for (int i = 0; i < count; ++i) {
    if (cfgs[i].path != GSK_PATH_FUSED32) return bail(fail(GSK_ERR_UNSUPPORTED, "ensembles run the FP32 fused path"));
    gsk_sim* s = nullptr;
    const int rc = gsk_create(&cfgs[i], &s);
    if (rc != GSK_OK) return bail(rc);
    e->m.push_back(s);
  }
  *out = e;
  return GSK_OK;
}

int gsk_ensemble_destroy(gsk_ensemble* e) {
  if (!e) return GSK_OK;
  (void)hipSetDevice(e->dev);
  if (e->master) (void)hipStreamSynchronize(e->master);
  ens_drop_graph(e);
  ens_drop_classes(e);
  for (gsk_sim* s : e->m) gsk_destroy(s);
  for (float* p : e->sharedB) (void)hipFree(p);
  for (hipStream_t st : e->bs) (void)hipStreamDestroy(st);
  for (hipEvent_t ev : e->join) (void)hipEventDestroy(ev);
  if (e->fork) (void)hipEventDestroy(e->fork);
  if (e->master) (void)hipStreamDestroy(e->master);
  delete e;
  return GSK_OK;
}

int gsk_ensemble_size(const gsk_ensemble* e) { return e ? (int)e->m.size() : -1; }

int gsk_ensemble_member(gsk_ensemble* e, int i, gsk_sim** out) {
  if (!e || !out || i < 0 || i >= (int)e->m.size()) return fail(GSK_ERR_ARG, "bad member index");
  *out = e->m[i];
  return GSK_OK;
}

int gsk_ensemble_build_btensor(gsk_ensemble* e, const double* C81, const double* eps, int neps) {
  if (!e || !C81 || !eps) return fail(GSK_ERR_ARG, "null argument");
  // one tensor per (grid, variant count, physics, spacing, wave-number convention)
  std::vector<gsk_sim*> owner;
  for (gsk_sim* s : e->m) {
    gsk_sim* o = nullptr;
    for (gsk_sim* c : owner)
      if (c->n == s->n && c->V == s->V && c->cfg.physics == s->cfg.physics && c->cfg.dx == s->cfg.dx &&
          c->cfg.kconv == s->cfg.kconv && c->g.W == s->g.W)
        o = c;
    if (!o) {
      CHK(gsk_build_btensor(s, C81, eps, neps));
      if (!s->sharedB) {  // the ensemble takes ownership of this member's tensor
        e->sharedB.push_back(s->Bt);
        s->sharedB = true;
      }
      owner.push_back(s);
      continue;
    }
    if (s->Bt && !s->sharedB) (void)hipFree(s->Bt);
    s->Bt = o->Bt;
    s->sharedB = true;
    s->cee0 = o->cee0;
    s->bargs = o->bargs;
    s->haveBargs = o->haveBargs;
    s->haveB = true;
    s->dirty = true;
    ++s->gen;
  }
  return GSK_OK;
}

int gsk_ensemble_set_schedule(gsk_ensemble* e, int mode, int branches, int steps_per_graph) {
  if (!e) return fail(GSK_ERR_ARG, "null handle");
  if (mode < GSK_ENS_SEQUENTIAL || mode > GSK_ENS_BATCHED_GRAPH) return fail(GSK_ERR_ARG, "unknown schedule");
  if (branches < 0 || steps_per_graph < 1) return fail(GSK_ERR_ARG, "bad schedule parameters");
  e->mode = mode;
  e->branches = branches;
  e->spg = steps_per_graph;
  return GSK_OK;
}

int gsk_ensemble_get_schedule(const gsk_ensemble* e, int* mode, int* branches, int* steps_per_graph) {
  if (!e || !mode || !branches || !steps_per_graph) return fail(GSK_ERR_ARG, "null argument");
  *mode = e->mode;
  *branches = ens_branches(e);
  *steps_per_graph = e->spg;
  return GSK_OK;
}

int gsk_ensemble_step(gsk_ensemble* e, int nsteps) {
  if (!e) return fail(GSK_ERR_ARG, "null handle");
  if (nsteps < 0) return fail(GSK_ERR_ARG, "negative step count");
  return ens_run(e, nsteps);
}

int gsk_ensemble_autotune(gsk_ensemble* e, int trial_steps, double* table, int max_rows, int* rows) {
  if (!e || trial_steps < 1) return fail(GSK_ERR_ARG, "bad argument");
  CHK(ens_prepare(e));
  EnsSnapshot snap;
  CHK(ens_save(e, snap));
  struct Cand { int mode, branches, spg; double ms; };
  std::vector<Cand> cand;
  const int cnt = (int)e->m.size();
  cand.push_back({GSK_ENS_SEQUENTIAL, 1, 1, 0.0});
  for (int b : {1, 2, 4, 8, 16, 32})
    if (b <= cnt) {
      cand.push_back({GSK_ENS_STREAMS, b, 1, 0.0});
      for (int k : {1, 5, 10})
        if (k <= trial_steps) cand.push_back({GSK_ENS_GRAPH, b, k, 0.0});
    }
  // batched launches per small-grid size class; the branches serve the other members
  for (int b : {8, 16, 32})
    if (b <= cnt || b == 8) {
      const int bb = std::min(b, cnt);
      cand.push_back({GSK_ENS_BATCHED, bb, 1, 0.0});
      for (int k : {1, 5, 10})
        if (k <= trial_steps) cand.push_back({GSK_ENS_BATCHED_GRAPH, bb, k, 0.0});
    }
  const int mode0 = e->mode, b0 = e->branches, k0 = e->spg;
  int rc = GSK_OK;
  for (Cand& c : cand) {
    e->mode = c.mode;
    e->branches = c.branches;
    e->spg = c.spg;
    rc = ens_run(e, c.spg);  // warm-up (and graph build)
    const auto t0 = std::chrono::steady_clock::now();
    if (rc == GSK_OK) rc = ens_run(e, trial_steps);
    const auto t1 = std::chrono::steady_clock::now();
    if (rc != GSK_OK) break;
    c.ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / trial_steps;
  }
  const int rr = ens_restore(e, snap);
  if (rc != GSK_OK || rr != GSK_OK) {
    e->mode = mode0; e->branches = b0; e->spg = k0;
    return rc != GSK_OK ? rc : rr;
  }
  const Cand* best = &cand[0];
  for (const Cand& c : cand)
    if (c.ms < best->ms) best = &c;
  e->mode = best->mode;
  e->branches = best->branches;
  e->spg = best->spg;
  if (rows) *rows = (int)cand.size();
  if (table)
    for (int i = 0; i < (int)cand.size() && i < max_rows; ++i) {
      table[4 * i + 0] = cand[i].mode;
      table[4 * i + 1] = cand[i].branches;
      table[4 * i + 2] = cand[i].spg;
      table[4 * i + 3] = cand[i].ms;
    }
  return GSK_OK;
}

}  // extern "C"

// ======================================================================== adjoint
// Reverse mode of the FP32 martensite step with checkpointing. gsk_adjoint_forward steps
// the handle from its current state and keeps (real field, k-space state, step) every K
// steps plus the final state; gsk_adjoint_backward walks the segments backwards: it
// restores a checkpoint, recomputes the segment's K states into a buffer (the restored
// state reproduces the original trajectory bit for bit), and runs the reverse steps over
// it. K = 1 keeps every state and recomputes nothing (store-all).
struct AdjWS {
  int T = 0, K = 0;
  long long step_end = 0;
  std::vector<float*> ck_fld;
  std::vector<float2*> ck_state;
  std::vector<long long> ck_step;
  float* fin_fld = nullptr;
  float2* fin_state = nullptr;
  float *seg = nullptr, *lam = nullptr, *mu = nullptr, *rq = nullptr, *bh = nullptr;
  float2 *kb = nullptr, *k2 = nullptr;
  double* part = nullptr;
  std::vector<double> hpart;
  int nb = 512;
  double ms_fwd = 0.0, ms_bwd = 0.0;
  double bytes_ck = 0.0, bytes_seg = 0.0, bytes_work = 0.0;
};

namespace {

void adj_free(AdjWS* w) {
  if (!w) return;
  for (float* p : w->ck_fld) (void)hipFree(p);
  for (float2* p : w->ck_state) (void)hipFree(p);
  free_dev(w->fin_fld); free_dev(w->fin_state);
  free_dev(w->seg); free_dev(w->lam); free_dev(w->mu); free_dev(w->rq); free_dev(w->bh);
  free_dev(w->kb); free_dev(w->k2); free_dev(w->part);
  delete w;
}

// copy the evolving state (real fields, k-space state) of the handle into / out of a slot
int adj_put(gsk_sim* s, float* fld, float2* st) {
  HIPCHK(hipMemcpyAsync(fld, s->fld, (size_t)s->F * nvox(s) * sizeof(float), hipMemcpyDeviceToDevice, s->stream));
  HIPCHK(hipMemcpyAsync(st, s->state, state_count(s) * sizeof(float2), hipMemcpyDeviceToDevice, s->stream));
  return GSK_OK;
}
int adj_get(gsk_sim* s, const float* fld, const float2* st, long long step) {
  HIPCHK(hipMemcpyAsync(s->fld, fld, (size_t)s->F * nvox(s) * sizeof(float), hipMemcpyDeviceToDevice, s->stream));
  HIPCHK(hipMemcpyAsync(s->state, st, state_count(s) * sizeof(float2), hipMemcpyDeviceToDevice, s->stream));
  const uint32_t v = (uint32_t)step;
  HIPCHK(hipMemcpyAsync(s->dstep, &v, sizeof(v), hipMemcpyHostToDevice, s->stream));
  CHK(sync(s));
  s->step = step;
  s->sq_ready = false;  // new fields: the next step transforms eta^2 afresh
  return GSK_OK;
}

template <typename T>
int dev_alloc(T*& p, size_t count, double& tally) {
  if (hipMalloc((void**)&p, count * sizeof(T)) != hipSuccess)
    return fail(GSK_ERR_HIP, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
  tally += (double)count * sizeof(T);
  return GSK_OK;
}

AdjK adj_consts(const gsk_sim* s, uint32_t step) {
  const MartK m = mart_consts(s);
  AdjK k{};
  k.A = m.A; k.B = m.B; k.C = m.C; k.two_wel = m.two_wel; k.noise = m.noise;
  k.dtL = m.dtL; k.dtLk = m.dtLk; k.kfac = m.kfac; k.inv_n3 = m.inv_n3;
  k.V = s->V;
  k.step = step;
  k.seed = m.seed;
  return k;
}

}  // namespace

extern "C" {

int gsk_adjoint_forward(gsk_sim* s, int nsteps, int K) {
  if (!s || nsteps < 1 || K < 1) return fail(GSK_ERR_ARG, "bad argument");
  if (s->cfg.path != GSK_PATH_FUSED32 || s->cfg.physics != GSK_PHYS_MART || s->cfg.rule != GSK_RULE_SPECTRAL)
    return fail(GSK_ERR_UNSUPPORTED, "the adjoint covers the FP32 fused spectral martensite step");
  if (s->V > 4) return fail(GSK_ERR_UNSUPPORTED, "the adjoint supports up to 4 variants");
  CHK(ensure_device(s->dev));
  CHK(prepare(s));
  adj_free(s->adj);
  s->adj = new AdjWS();
  AdjWS& w = *s->adj;
  w.T = nsteps;
  w.K = K;
  const int nck = (nsteps + K - 1) / K;
  const size_t fcount = (size_t)s->F * nvox(s), scount = state_count(s);
  const auto t0 = std::chrono::steady_clock::now();
  for (int j = 0; j < nck; ++j) {
    float* f = nullptr;
    float2* t = nullptr;
    CHK(dev_alloc(f, fcount, w.bytes_ck));
    w.ck_fld.push_back(f);
    CHK(dev_alloc(t, scount, w.bytes_ck));
    w.ck_state.push_back(t);
    w.ck_step.push_back(s->step);
    CHK(adj_put(s, f, t));
    const int len = std::min(K, nsteps - j * K);
    for (int i = 0; i < len; ++i) {
      CHK(dispatch_n<DStep>(s->n, s));
      ++s->step;
    }
  }
  CHK(dev_alloc(w.fin_fld, fcount, w.bytes_ck));
  CHK(dev_alloc(w.fin_state, scount, w.bytes_ck));
  CHK(adj_put(s, w.fin_fld, w.fin_state));
  CHK(sync(s));
  w.step_end = s->step;
  w.ms_fwd = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return GSK_OK;
}

int gsk_adjoint_backward(gsk_sim* s, const float* dL_detaT, float* grad_eta0, double* grad_params) {
  if (!s || !dL_detaT) return fail(GSK_ERR_ARG, "null argument");
  if (!s->adj || s->adj->T == 0) return fail(GSK_ERR_STATE, "no recorded forward run (call gsk_adjoint_forward)");
  if (!s->haveBargs) return fail(GSK_ERR_STATE, "B tensor not built");
  CHK(ensure_device(s->dev));
  AdjWS& w = *s->adj;
  const int V = s->V, K = w.K, T = w.T, nck = (T + K - 1) / K, nb = w.nb;
  const size_t N3 = nvox(s), VN = (size_t)V * N3;
  if (!w.lam) {
    if (K > 1) CHK(dev_alloc(w.seg, (size_t)(K + 1) * VN, w.bytes_seg));
    CHK(dev_alloc(w.lam, VN, w.bytes_work));
    CHK(dev_alloc(w.mu, VN, w.bytes_work));
    CHK(dev_alloc(w.rq, 2 * VN, w.bytes_work));
    CHK(dev_alloc(w.kb, (size_t)(2 * V) * s->g.kfs, w.bytes_work));
    CHK(dev_alloc(w.k2, (size_t)V * s->g.kfs, w.bytes_work));
    CHK(dev_alloc(w.bh, (size_t)(V * (V + 1) / 2) * s->g.kfs, w.bytes_work));
    // FP64 partial sums of every reverse step of a segment, copied to the host once per segment
    CHK(dev_alloc(w.part, (size_t)K * nb * (2 + V * V), w.bytes_work));
    w.hpart.resize((size_t)K * nb * (2 + V * V));
    hipLaunchKernelGGL(k_btensor_half, dim3(grid_1d((long)s->n * s->n * (s->n / 2 + 1))), dim3(256), 0, s->stream,
                       w.bh, s->g, s->bargs);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipMemcpyAsync(w.lam, dL_detaT, VN * sizeof(float), hipMemcpyHostToDevice, s->stream));
  const double dt = s->cfg.dt, dtL = s->cfg.dt * s->cfg.L, kap = s->cfg.kappa, wel = s->cfg.w_el;
  const double inv_n3 = 1.0 / (double)N3;
  double gL = 0.0, gk = 0.0;
  std::vector<double> gs(V, 0.0);
  const auto t0 = std::chrono::steady_clock::now();
  for (int j = nck - 1; j >= 0; --j) {
    const int len = std::min(K, T - j * K);
    CHK(adj_get(s, w.ck_fld[j], w.ck_state[j], w.ck_step[j]));
    if (K > 1) {  // recompute the segment's states
      HIPCHK(hipMemcpyAsync(w.seg, s->fld, VN * sizeof(float), hipMemcpyDeviceToDevice, s->stream));
      for (int i = 1; i <= len; ++i) {
        CHK(dispatch_n<DStep>(s->n, s));
        ++s->step;
        HIPCHK(hipMemcpyAsync(w.seg + (size_t)i * VN, s->fld, VN * sizeof(float), hipMemcpyDeviceToDevice, s->stream));
      }
    }
    const size_t stride = (size_t)nb * (2 + V * V);
    for (int i = len - 1; i >= 0; --i) {
      const float* eta = K > 1 ? w.seg + (size_t)i * VN : w.ck_fld[j];
      const float* eta1 = K > 1 ? w.seg + (size_t)(i + 1) * VN : (j + 1 < nck ? w.ck_fld[j + 1] : w.fin_fld);
      const AdjK k = adj_consts(s, (uint32_t)(w.ck_step[j] + i));
      CHK(dispatch_n<DAdj>(s->n, s, eta, eta1, w.lam, w.mu, w.rq, w.kb, w.k2, (const float*)w.bh,
                           w.part + (size_t)i * stride, nb, k));
    }
    HIPCHK(hipMemcpyAsync(w.hpart.data(), w.part, (size_t)len * stride * sizeof(double), hipMemcpyDeviceToHost,
                          s->stream));
    CHK(sync(s));
    // the step's sums in the same order as the reverse sweep: gradients do not depend on K
    for (int i = len - 1; i >= 0; --i) {
      const double* hp = w.hpart.data() + (size_t)i * stride;
      double X = 0.0, muF = 0.0;
      std::vector<double> G((size_t)V * V, 0.0);
      for (int b = 0; b < nb; ++b) {
        X += hp[b];
        for (int r = 0; r < V * V; ++r) G[r] += hp[nb + (size_t)b * V * V + r];
        muF += hp[(size_t)nb + (size_t)nb * V * V + b];
      }
      X *= inv_n3;
      gL += -dt * kap * X - dt * muF;
      gk += -dtL * X;
      for (int r = 0; r < V; ++r)
        for (int q = 0; q < V; ++q) {
          const double t = -2.0 * dtL * wel * G[r * V + q] * inv_n3;
          gs[r] += t;
          gs[q] += t;
        }
    }
  }
  if (grad_eta0) HIPCHK(hipMemcpyAsync(grad_eta0, w.lam, VN * sizeof(float), hipMemcpyDeviceToHost, s->stream));
  CHK(adj_get(s, w.fin_fld, w.fin_state, w.step_end));
  w.ms_bwd = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (grad_params) {
    for (int v = 0; v < V; ++v) grad_params[v] = gs[v];
    grad_params[V] = gL;
    grad_params[V + 1] = gk;
  }
  return GSK_OK;
}

int gsk_adjoint_info(const gsk_sim* s, double* out) {
  if (!s || !out) return fail(GSK_ERR_ARG, "null argument");
  if (!s->adj) return fail(GSK_ERR_STATE, "no recorded forward run");
  const AdjWS& w = *s->adj;
  out[0] = w.bytes_ck;
  out[1] = w.bytes_seg;
  out[2] = w.bytes_work;
  out[3] = w.ms_fwd;
  out[4] = w.ms_bwd;
  out[5] = (double)w.ck_fld.size() + 1.0;  // checkpoints incl. the final state
  return GSK_OK;
}

int gsk_adjoint_release(gsk_sim* s) {
  if (!s) return fail(GSK_ERR_ARG, "null handle");
  CHK(ensure_device(s->dev));
  adj_free(s->adj);
  s->adj = nullptr;
  return GSK_OK;
}

}  // extern "C"

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
    hipLaunchKernelGGL(k_btensor_half_rect, dim3(grid_1d((long)n * n * (n / 2 + 1))), dim3(256), 0, s->stream, bh,
                       s->g, a);
    if (source == 2)
      hipLaunchKernelGGL(k_square, dim3(grid_1d((long)nvar * N3)), dim3(256), 0, s->stream, s->fld, (long)nvar * N3);
    if (hipGetLastError() != hipSuccess) rc = fail(GSK_ERR_HIP, "launch failed");
  }
  step(rc == GSK_OK ? dispatch_n<DR2C>(n, s, (const float*)s->fld, s->kbuf, nvar) : rc);
  if (rc == GSK_OK) {
    hipLaunchKernelGGL(k_contract_rect, dim3(grid_1d((long)n * n * (n / 2 + 1))), dim3(256), 0, s->stream, s->kbuf,
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
