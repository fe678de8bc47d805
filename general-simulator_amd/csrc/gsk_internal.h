// Internal declarations of the library (libgsk.so) shared by its translation units:
//   gsk_api.hip       C ABI: handles, fields, steps, reductions, timing, noise, B tensor, FFTs
//   gsk_impl.hip      the per-size step schedules and kernel launches (one object per grid edge)
//   gsk_ensemble.hip  ensembles: schedules, batched launches, hipGraph capture, autotuner
//   gsk_adjoint.hip   checkpointed reverse mode
//   gsk_stress.hip    the stress plugin's transform
//   fft_compare.hip   the rocFFT comparator
#pragma once
#include "gsk.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "kernels_adjoint.h"
#include "kernels_aux.h"
#include "kernels_plane.h"
#include "kernels_split.h"
#include "kernels_ref.h"

using namespace gsk;


// the message of the last failure on this thread (defined once, in gsk_api.hip)
std::string& err_text();

inline int fail(int code, const std::string& msg) {
  err_text() = msg;
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

inline bool supported_n(int n) {
  switch (n) {
    case 16: case 32: case 64: case 96: case 128: case 192: case 256: case 512: return true;
    default: return false;
  }
}

inline int rmin_of(int N) {
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

inline int round64(int t) { return ((t + 63) / 64) * 64; }

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
inline int xupd_group(int left) {
  if (GSK_XUPD_G == 0) return left;  // a single launch covers every field
  return (GSK_XUPD_G >= 3 && left >= 3) ? 3 : ((GSK_XUPD_G >= 2 && left >= 2) ? 2 : 1);
}

// block size of the z kernels for half-length M
inline int zthreads_host(int M) {
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

struct AdjWS;  // adjoint workspace (checkpoints and buffers), defined with the adjoint entry points

// Box grids (edges nx, ny, nz or spacings dx, dy, dz not all equal): the plain-path kernels on
// per-axis transforms (gsk_box.hip). Buffers hold T = float (GSK_PATH_FUSED32) or double
// (GSK_PATH_PLAIN64) values.
struct BoxWS {
  bool f64 = false;
  double dx = 1, dy = 1, dz = 1;
  void* fld = nullptr;                // F real fields [f][x][y][z]: the state
  void *S = nullptr, *R = nullptr;    // V complex fields: FFT(eta^2), the contraction R
  void *E = nullptr, *Fh = nullptr;   // F complex fields: the transformed fields and forces
  void* B = nullptr;                  // [p][q][voxel] (V x V) or [voxel] (the precipitate's c source)
  void* tw[3] = {nullptr, nullptr, nullptr};  // twiddles of the x, y and z transform lengths
  int L = 1, W[2] = {1, 1};           // z rows per tile; column chunk of the x and y passes
  int tz = 64, tx = 64, ty = 64;      // threads of the z, x and y passes
  size_t ldz = 0, ldx = 0, ldy = 0;   // their LDS bytes
};

struct gsk_sim {
  gsk_config cfg{};
  int n = 0, V = 0, F = 0, dev = 0;
  int nx = 0, ny = 0, nz = 0;  // grid edges (all n for a cube)
  BoxWS* bx = nullptr;         // box grid: its buffers (null for a cube)
  Geo g{};
  int zthreads = 0, xthreads = 0;
  int cus = 0;  // compute units of the device (0 until first needed)
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
  double* fstat = nullptr;    // per-block sums of the four force terms of the last reference step
  int fblocks = 0;            // work groups (partial sums in fstat) of the last reference step
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
  // split-y schedule (n = 256, spectral martensite): its geometry; the B tensor and the k-space
  // state are then in its tile order, and the carried S_hat is in E-form
  bool split = false;
  SGeo sg{};
  int order = 0;  // visiting order of the last per-field launch
  long long step = 0;
  // profiling
  bool prof = false;
  std::vector<hipEvent_t> ev;
  int evi = 0;
  std::vector<Slot> slots;
};


// Geometry for the next per-field launch: flips the field / tile visiting order each time.
inline Geo geo_next(gsk_sim* s) {
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
inline bool plane_on(const gsk_sim* s) {
  static const bool off = [] {
    const char* e = std::getenv("GSK_PLANE");
    return e && e[0] == '0';
  }();
  return !off && (s->n == 32 || s->n == 64 || s->n == 128) && s->F >= 1 && s->F <= 4 &&
         s->cfg.path == GSK_PATH_FUSED32 && s->cfg.physics == GSK_PHYS_MART && s->cfg.rule == GSK_RULE_SPECTRAL;
}

// Split-y schedule (kernels_split.h) at n = 128, 192, 256: the martensite step with 1..4 variants (the
// spectral rule's step; the reference rule's elastic R) and the precipitate step with 2..5 fields
// (both rules; the reference rule's elastic R of up to 4 order parameters). GSK_SPLIT=0 in the
// environment when a handle is created keeps that handle on the chain of y and z kernels.
inline bool split_wanted(const gsk_sim* s) {
  const char* e = std::getenv("GSK_SPLIT");
  const bool on = !(e && e[0] == '0');
  if (!on || (s->n != 256 && s->n != 192 && s->n != 128) || s->cfg.path != GSK_PATH_FUSED32) return false;
  if (s->cfg.physics == GSK_PHYS_MART) return s->F >= 1 && s->F <= 4;
  return s->F >= 2 && s->F <= 5;
}

inline size_t nvox(const gsk_sim* s) { return (size_t)s->nx * s->ny * s->nz; }
inline size_t hs_count(const gsk_sim* s) { return (size_t)s->n * s->n * (s->n / 2 + 1); }
// State element of field 0 at k = 0: first main x tile (after the n/W Nyquist tiles), kx = 0, column 0.
inline size_t k0_index(const gsk_sim* s) {
  if (s->split)  // the first main tile (k1 = 0, kz from 0) after the Nyquist tiles, column 0
    return (size_t)(s->sg.m / kSplitWc) * s->F * s->n * split_w<kSplitWc>();
  return (size_t)(s->n / s->g.W) * s->F * s->n * s->g.W;
}

inline int mark(gsk_sim* s) {
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

inline MartK mart_consts(const gsk_sim* s) {
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

inline PrecK prec_consts(const gsk_sim* s) {
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

inline PlainK plain_consts(const gsk_sim* s) {
  const gsk_config& c = s->cfg;
  PlainK k{};
  k.A = c.A; k.B = c.B; k.C = c.C;
  k.two_wel = 2.0 * c.w_el;
  k.noise = c.noise;
  k.dtL = c.dt * c.L;
  k.dtLk = c.dt * c.L * c.kappa;
  k.kfx = k.kfy = k.kfz = (c.kconv == GSK_K_REFERENCE ? 2.0 * 3.14 * c.dx / s->n : 2.0 * kPi / (s->n * c.dx));
  k.d = BoxDim{s->n, s->n, s->n};
  k.inv_n3 = 1.0 / ((double)s->n * s->n * s->n);
  k.V = s->V;
  k.step = (uint32_t)s->step;
  k.seed = c.seed;
  return k;
}

inline PlainPK plain_prec_consts(const gsk_sim* s) {
  const gsk_config& c = s->cfg;
  PlainPK k{};
  for (int i = 0; i < 8; ++i) k.A[i] = c.a[i];
  k.c1 = c.c1; k.c2 = c.c2;
  k.dtM = c.dt * c.M;
  k.dtMkc = c.dt * c.M * c.kappa_c;
  k.dtL = c.dt * c.L;
  k.dtLke = c.dt * c.L * c.kappa;
  k.xi = c.xi;
  k.S = c.S;
  k.noise = c.noise;
  k.kfx = k.kfy = k.kfz = (c.kconv == GSK_K_REFERENCE ? 2.0 * 3.14 * c.dx / s->n : 2.0 * kPi / (s->n * c.dx));
  k.d = BoxDim{s->n, s->n, s->n};
  k.celastic = 1;
  k.inv_n3 = 1.0 / ((double)s->n * s->n * s->n);
  k.V = s->V;
  k.step = (uint32_t)s->step;
  k.seed = c.seed;
  return k;
}

// blocks of the martensite reference-rule kernel (its per-block force sums: kRefBlocks * 4 doubles)
constexpr int kRefBlocks = 2048;

inline int grid_1d(long work, int threads = 256) {
  long b = (work + threads - 1) / threads;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (int)b;
}

// The martensite reference-rule update of the V fields at eta on an nx x ny x nz grid, R in RE on
// entry (the new eta on exit): the 16-byte XCD-ranged kernel for 1..4 variants, the general one
// otherwise. Records the work-group count whose force sums fstat holds.
inline int launch_ref_mart(gsk_sim* s, const float* eta, float* RE, int nx, int ny, int nz, const RefMartK& k) {
  const long nq = (long)nx * ny * nz / 4;
  const int V = s->V;
  if (V < 1 || V > 4) {
    s->fblocks = (int)std::min<long>(grid_1d(nq), kRefBlocks);
    LAUNCH(k_ref_mart<>, s->fblocks, 256, 0, s, eta, RE, s->defect, s->Kd, nx, ny, nz, k, s->fstat);
    return GSK_OK;
  }
  if (!s->cus) HIPCHK(hipDeviceGetAttribute(&s->cus, hipDeviceAttributeMultiprocessorCount, s->dev));
  const long chunks = (nq + 255) / 256;  // 256 quads per work group and chunk
  const long g = std::min<long>(std::min<long>(kRefBlocks, 4L * s->cus), (chunks + 7) / 8 * 8);
  const int G = (int)std::max<long>(8, g / 8 * 8);
  const int L = nz / 4;
  const bool zs = L <= 64 && 64 % L == 0 && nq % 256 == 0;
  s->fblocks = G;
#define GSK_REFQ(v, z) LAUNCH((k_ref_mart_q<v, z>), G, 256, 0, s, eta, RE, s->defect, s->Kd, nx, ny, nz, k, s->fstat)
  switch (V * 2 + (zs ? 1 : 0)) {
    case 2: GSK_REFQ(1, false); break;
    case 3: GSK_REFQ(1, true); break;
    case 4: GSK_REFQ(2, false); break;
    case 5: GSK_REFQ(2, true); break;
    case 6: GSK_REFQ(3, false); break;
    case 7: GSK_REFQ(3, true); break;
    case 8: GSK_REFQ(4, false); break;
    default: GSK_REFQ(4, true);
  }
#undef GSK_REFQ
  return GSK_OK;
}

// Allow a kernel the device's whole LDS for dynamic shared memory (minus its static part).
inline int lds_limit(const void* f) {
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
inline bool prec_carries(const gsk_sim* s) {
  return s->cfg.path == GSK_PATH_FUSED32 && s->cfg.physics == GSK_PHYS_PREC && s->cfg.rule == GSK_RULE_SPECTRAL &&
         s->F >= 2 && s->F <= 5;
}
inline bool carries(const gsk_sim* s) {
  return prec_carries(s) || (s->cfg.path == GSK_PATH_FUSED32 && s->cfg.physics == GSK_PHYS_MART &&
                             s->cfg.rule == GSK_RULE_SPECTRAL);
}
// Entry points of the per-size implementation (gsk_impl.hip, compiled once per grid edge N with
// -DGSK_IMPL_N=N; each object file holds the explicit instantiations of its N).
template <int N> int impl_set_attrs();
template <int N> int impl_init_state(gsk_sim* s);
template <int N> int impl_square_fwd(gsk_sim* s);
template <int N> int impl_step_batched(gsk_sim* s0, int cnt, const MemberPtrs* tab, hipStream_t st);
template <int N> int impl_step(gsk_sim* s);
template <int N> int impl_energy_el_fused(gsk_sim* s, double* el);
template <int N> int impl_energy_el_plain(gsk_sim* s, double* gr, double* el);
template <int N> int impl_energy_prec_plain(gsk_sim* s, double* out);
template <int N> int impl_zfwd(gsk_sim* s, const float* in, int F);
template <int N> int impl_ypass(gsk_sim* s, bool fwd, int F);
template <int N> int impl_xpass(gsk_sim* s, bool fwd, int F, float scale);
template <int N> int impl_adj_step(gsk_sim* s, const float* eta, const float* eta1, float* lam, float* mu, float* rq,
                                   float2* kb, float2* k2, const float* bh, double* part, int nb, const AdjK& k);
template <int N> int impl_r2c(gsk_sim* s, const float* in, float2* kb, int F);
template <int N> int impl_c2r(gsk_sim* s, float2* kb, float* out, int F);
template <int N> int impl_zinv(gsk_sim* s, float* out, int F);

// Box grids (gsk_box.hip)
bool box_grid(const gsk_config& c);  // edges or spacings not all equal
int box_create(gsk_sim* s);
void box_free(BoxWS* w);
int box_build_btensor(gsk_sim* s, const BArgs& a);
int box_set_field(gsk_sim* s, int off_fields, size_t count, const double* host);
int box_get_field(gsk_sim* s, int off_fields, size_t count, double* host);
int box_step(gsk_sim* s);
int box_energy(gsk_sim* s, double* gr, double* el);

template <template <int> class Fn, typename... Args>
inline int dispatch_n(int n, Args&&... args) {
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

template <int N> struct DAttrs { static int run() { return impl_set_attrs<N>(); } };
template <int N> struct DInit { static int run(gsk_sim* s) { return impl_init_state<N>(s); } };
template <int N> struct DSquare { static int run(gsk_sim* s) { return impl_square_fwd<N>(s); } };
template <int N> struct DBatch {
  static int run(gsk_sim* s0, int cnt, const MemberPtrs* tab, hipStream_t st) {
    return impl_step_batched<N>(s0, cnt, tab, st);
  }
};
template <int N> struct DStep { static int run(gsk_sim* s) { return impl_step<N>(s); } };
template <int N> struct DElF { static int run(gsk_sim* s, double* el) { return impl_energy_el_fused<N>(s, el); } };
template <int N> struct DElP {
  static int run(gsk_sim* s, double* gr, double* el) { return impl_energy_el_plain<N>(s, gr, el); }
};
template <int N> struct DEpP { static int run(gsk_sim* s, double* out) { return impl_energy_prec_plain<N>(s, out); } };
template <int N> struct DZfwd { static int run(gsk_sim* s, const float* in, int F) { return impl_zfwd<N>(s, in, F); } };
template <int N> struct DY { static int run(gsk_sim* s, bool fwd, int F) { return impl_ypass<N>(s, fwd, F); } };
template <int N> struct DX {
  static int run(gsk_sim* s, bool fwd, int F, float scale) { return impl_xpass<N>(s, fwd, F, scale); }
};
template <int N> struct DAdj {
  static int run(gsk_sim* s, const float* eta, const float* eta1, float* lam, float* mu, float* rq, float2* kb,
                 float2* k2, const float* bh, double* part, int nb, const AdjK& k) {
    return impl_adj_step<N>(s, eta, eta1, lam, mu, rq, kb, k2, bh, part, nb, k);
  }
};
template <int N> struct DR2C {
  static int run(gsk_sim* s, const float* in, float2* kb, int F) { return impl_r2c<N>(s, in, kb, F); }
};
template <int N> struct DC2R {
  static int run(gsk_sim* s, float2* kb, float* out, int F) { return impl_c2r<N>(s, kb, out, F); }
};
template <int N> struct DZinv { static int run(gsk_sim* s, float* out, int F) { return impl_zinv<N>(s, out, F); } };

// one step of a handle: the per-size schedule of a cube, or the box path
inline int step_once(gsk_sim* s) { return s->bx ? box_step(s) : dispatch_n<DStep>(s->n, s); }

inline int ensure_device(int dev) {
  int cnt = 0;
  HIPCHK(hipGetDeviceCount(&cnt));
  if (cnt <= 0) return fail(GSK_ERR_HIP, "no HIP device available");
  if (dev < 0 || dev >= cnt) return fail(GSK_ERR_ARG, "device ordinal out of range");
  HIPCHK(hipSetDevice(dev));
  return GSK_OK;
}

// Choose tile shapes for a given field count F.
inline void choose_geometry(gsk_sim* s, int Fz, int Fx_resident) {
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
  s->split = split_wanted(s);
  if (s->split) {
    SGeo& q = s->sg;
    q.n = n; q.M = M; q.m = n / kSplitP;
    q.ntile = split_ntile<kSplitWc>(q.m, M);
    q.ntile_c = split_ntile<kSplitWcC>(q.m, M);
    q.kfs = g.kfs; q.nq = g.nq; q.nvox = g.nvox;
  }
}

inline void choose_plain(gsk_sim* s) {
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

inline void build_slots(gsk_sim* s) {
  s->slots.clear();
  const double hs = (double)hs_count(s) * 8.0;  // one half-spectrum complex field
  const double rv = (double)nvox(s) * 4.0;      // one real fp32 field
  const int F = s->F;
  if (s->cfg.path == GSK_PATH_PLAIN64) {
    // one kernel per operation; real fields share C2C transforms (pairs, or a force with its own
    // field), and the unpacking kernels read each packed spectrum at k and at -k
    const double cf = (double)nvox(s) * 16.0, r8 = (double)nvox(s) * 8.0;
    auto add_c2c = [&](int nf, const char* a, const char* b, const char* c) {
      for (int f = 0; f < nf; ++f) {
        s->slots.push_back({a, 2 * cf});
        s->slots.push_back({b, 2 * cf});
        s->slots.push_back({c, 2 * cf});
      }
    };
    if (s->cfg.physics == GSK_PHYS_PREC) {
      const int PF = (F + 1) / 2;
      s->slots.push_back({"p_prec_local", F * (r8 + cf)});
      add_c2c(F, "p_z_fwdG", "p_y_fwdG", "p_x_fwdG");
      s->slots.push_back({"p_prec_update_pair", 2 * F * cf + r8 + PF * cf});
      add_c2c(PF, "p_z_invE", "p_y_invE", "p_x_invE");
      s->slots.push_back({"p_extract_pair", PF * cf + F * r8});
      return;
    }
    const int V = s->V, P = (V + 1) / 2;
    s->slots.push_back({"p_square_pair", V * r8 + P * cf});
    add_c2c(P, "p_z_fwdS", "p_y_fwdS", "p_x_fwdS");
    s->slots.push_back({"p_contract_pair", 3 * P * cf + V * V * r8});
    add_c2c(P, "p_z_invR", "p_y_invR", "p_x_invR");
    s->slots.push_back({"p_force_pair", V * r8 + P * cf + V * cf});
    add_c2c(V, "p_z_fwdG", "p_y_fwdG", "p_x_fwdG");
    s->slots.push_back({"p_update_pair", 2 * V * cf + P * cf});
    add_c2c(P, "p_z_invE", "p_y_invE", "p_x_invE");
    s->slots.push_back({"p_extract_pair", P * cf + V * r8});
    return;
  }
  if (s->cfg.physics == GSK_PHYS_MART && s->cfg.rule == GSK_RULE_REFERENCE) {
    // elastic R in real space, then the explicit update (eta and R read, the new eta written
    // over R; the dislocation term K and the defect mask when set)
    const int npq = F * (F + 1) / 2;
    if (s->split) {  // three launches for R: sub-plane forward, contraction, sub-plane inverse
      s->slots.push_back({"sub_fwd_sq", F * (rv + hs)});
      s->slots.push_back({"x_contract", F * 2 * hs + npq * hs / 2});
      s->slots.push_back({"sub_inv_R", F * (hs + rv)});
    } else {
      s->slots.push_back({"zfwd_sq", F * (rv + hs)});
      s->slots.push_back({"y_fwd_S", F * 2 * hs});
      s->slots.push_back({"x_contract", F * 2 * hs + npq * hs / 2});
      s->slots.push_back({"y_inv_R", F * 2 * hs});
      s->slots.push_back({"zinv_R", F * (hs + rv)});
    }
    s->slots.push_back({"ref_update", F * 3 * rv + (s->Kd ? F * rv : 0.0) + (s->defect ? rv : 0.0)});
    s->slots.push_back({"advance", 0.0});
    return;
  }
  if (s->cfg.physics == GSK_PHYS_MART) {
    const int npq = F * (F + 1) / 2;
    if (s->split) {  // four launches: the split-y schedule
      s->slots.push_back({"x_contract", F * 2 * hs + npq * hs / 2});
      s->slots.push_back({"sub_mid", F * (2 * hs + rv)});  // R and F spectra, each eta once
      s->slots.push_back({"x_update_g0", F * 4 * hs});
      s->slots.push_back({"sub_sq", F * (2 * hs + rv)});  // eta and S spectra, eta written
      return;
    }
    const bool pl = plane_on(s);
    if (!pl) s->slots.push_back({"y_fwd_S", F * 2 * hs});
    s->slots.push_back({"x_contract", F * 2 * hs + npq * hs / 2});
    if (pl) {  // R and F spectra, each eta once (the variant groups share them through L2)
      s->slots.push_back({"plane_mid", F * (2 * hs + rv)});
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
  } else if (s->split) {  // precipitate, split-y schedule
    const bool ref = s->cfg.rule == GSK_RULE_REFERENCE;
    if (ref) {  // the elastic R of the V order parameters, then the local forces with it
      const int V = s->V, npq = V * (V + 1) / 2;
      s->slots.push_back({"sub_fwd_sq", V * (rv + hs)});
      s->slots.push_back({"x_contract", V * 2 * hs + npq * hs / 2});
      s->slots.push_back({"sub_inv_R", V * (hs + rv)});
      s->slots.push_back({"sub_local", F * (rv + hs) + V * rv});
    }
    s->slots.push_back({"x_update", F * 4 * hs + (ref ? 0.0 : hs / 2)});
    s->slots.push_back({"sub_inv", F * (hs + rv)});
    if (!ref) s->slots.push_back({"sub_local", F * (rv + hs)});  // each field read once (L2-shared)
  } else {
    if (s->cfg.rule == GSK_RULE_REFERENCE) {  // the elastic R of the V order parameters first
      const int V = s->V, npq = V * (V + 1) / 2;
      s->slots.push_back({"zfwd_sq", V * (rv + hs)});
      s->slots.push_back({"y_fwd_S", V * 2 * hs});
      s->slots.push_back({"x_contract", V * 2 * hs + npq * hs / 2});
      s->slots.push_back({"y_inv_R", V * 2 * hs});
      s->slots.push_back({"zinv_R", V * (hs + rv)});
    }
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

inline int sync(gsk_sim* s) {
  HIPCHK(hipStreamSynchronize(s->stream));
  return GSK_OK;
}

inline int prepare(gsk_sim* s) {
  if (!s->haveB) return fail(GSK_ERR_STATE, "B tensor not built (call gsk_build_btensor first)");
  if (s->bx) return GSK_OK;  // box grids carry no k-space state
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
inline void free_dev(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}


// complex elements of the tile-ordered k-space state of a handle
inline size_t state_count(const gsk_sim* s) { return (size_t)s->g.ntile * s->F * s->n * s->g.W; }

// Shared by the B tensor builders of gsk_api.hip and gsk_stress.hip.
int fill_bargs(BArgs& a, const double* C81, const double* e1, int V1, const double* e2, int V2, double kfac);
// Frees an adjoint workspace (gsk_adjoint.hip).
void adj_free(AdjWS* w);
