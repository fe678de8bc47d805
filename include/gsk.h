/*
 * gsk.h — C ABI of the MI355X phase-field time-step library (libgsk.so).
 *
 * The library owns device memory, HIP streams and kernels; callers pass plain
 * host pointers and sizes. Every function returns an int status (GSK_OK = 0,
 * negative on error, mirroring Code_* in src/pub.h:17-22 of General-Simulator)
 * and records a message readable with gsk_last_error().
 *
 * Threading: a handle is used by one host thread at a time (the reference runs
 * one host thread on the default stream); distinct handles may be used from
 * distinct threads. Calls are synchronous at the boundary.
 *
 * Reference interfaces replaced (General-Simulator, file:line):
 *   gsk_create            Dynamics_mart::Initialize / Dynamics_precipitate::Initialize
 *                         (src/dynamics_mart.cu:14-173, src/dynamics_precipitate.cu:17-110)
 *   gsk_build_btensor     GTensorB::InitB (src/gtensorb.cu:55-108)
 *   gsk_set/get_field     Data<T>::HostToDevice / DeviceToHost (include/datamain.th:283-294)
 *   gsk_init_uniform      initial random order parameters (Random::InitRandom, src/random.cu:52-63)
 *   gsk_step              Dynamics_*::Calculate per step (src/dynamics_mart.cu:348-417,
 *                         src/dynamics_precipitate.cu:278-289) driven by GS::Run (src/gs.cu:201-209)
 *   gsk_reduce            Data<T>::TotalDevice used by Get() (include/datamain.th:782-806,
 *                         src/dynamics_mart.cu:443-456), here in FP64
 *   gsk_noise_field       Random::NewNormal_device (src/random.cu:79-82), Philox4x32-10 here
 *   gsk_destroy           ~Dynamics_* (cufftDestroy, src/dynamics_mart.cu:177-179)
 */
#ifndef GSK_H
#define GSK_H

#ifdef __cplusplus
extern "C" {
#endif

#define GSK_OK 0
#define GSK_ERR (-1)          /* generic failure (Code_ERR)             */
#define GSK_ERR_ARG (-2)      /* invalid argument                       */
#define GSK_ERR_HIP (-3)      /* HIP runtime error                      */
#define GSK_ERR_STATE (-4)    /* call out of order (e.g. no B tensor)   */
#define GSK_ERR_UNSUPPORTED (-5)

/* physics */
#define GSK_PHYS_MART 0   /* martensite: semi-implicit Allen-Cahn, Khachaturyan elasticity on eta^2 */
#define GSK_PHYS_PREC 1   /* precipitate: Cahn-Hilliard c + Allen-Cahn eta, c-sourced elasticity    */
/* execution path */
#define GSK_PATH_FUSED32 0  /* FP32 fused spectral kernels (R2C, 3 axis passes per transform) */
#define GSK_PATH_PLAIN64 1  /* FP64 plain path: one kernel per operation, C2C transforms       */
/* wave-vector convention */
#define GSK_K_EXACT 0       /* k = 2 pi m / (n dx)                                  */
#define GSK_K_REFERENCE 1   /* k = 2 * 3.14 * dx * m / n  (src/gtensorb.cu:9-12)     */
/* host data types */
#define GSK_F32 0
#define GSK_F64 1
/* reductions */
#define GSK_RED_ENERGY 1    /* out[0..3] = chemical, gradient, elastic, total */
#define GSK_RED_MASS 2      /* out[0] = sum of c (FP64 mean x N + k=0 state),
                               out[1] = FP64 sum of the stored FP32 real field */
#define GSK_RED_MEAN_ETA 3  /* out[v] = mean of eta_v, v < nvar              */
#define GSK_RED_SUM_ETA2 4  /* out[0] = sum over voxels and variants of eta^2 */
#define GSK_RED_MART_FORCES 5 /* martensite reference rule: out[0..3] = means over voxels and variants of
                                 the gradient (w_grad Grad), chemical, elastic and dislocation force
                                 terms of the last step (zero before the first step); the reference's
                                 Get columns gradient/chemical/elastic/dislocation
                                 (src/dynamics_mart.cu:443-456) */

typedef struct gsk_sim gsk_sim;

typedef struct gsk_config {
  int n;                 /* grid edge nx (periodic): 16, 32, 64, 96, 128, 192, 256 or 512 */
  int nvar;              /* number of order parameters eta_v                              */
  int physics;           /* GSK_PHYS_*                                                    */
  int path;              /* GSK_PATH_*                                                    */
  int kconv;             /* GSK_K_*                                                       */
  int device;            /* HIP device ordinal                                            */
  double dx;             /* grid spacing                                                  */
  unsigned long long seed; /* Philox key for the noise                                    */
  /* Allen-Cahn (both physics) */
  double L, kappa, dt, noise;
  /* martensite Landau f = A/2 sum eta^2 - B/3 sum eta^3 + C/4 (sum eta^2)^2 */
  double A, B, C;
  double w_el;           /* elastic weight                                                */
  /* precipitate (src/dynamics_precipitate.cu:20-29): a[1..7], c1, c2, M, kappa_c, xi */
  double a[8];
  double c1, c2, M, kappa_c, xi;
  /* Cahn-Hilliard stabilisation S: the term S k^2 c_hat is moved to the implicit
   * side of the concentration update; S = 0 is the reference rule
   * (src/dynamics_precipitate.cu:203-205). */
  double S;
  /* Rule. GSK_RULE_SPECTRAL: the semi-implicit spectral rules above. GSK_RULE_REFERENCE:
   * the plugins' own updates (FP32 fused path only):
   *   martensite (src/dynamics_mart.cu:181-417): eta += dt (w_grad Grad + Chem + El + Dis
   *     + noise xi), eta *= (1 - defect); Chem from a1 = 0.02 (temperature - T0), a2 = B,
   *     a3 = C (the reference's LPC[1..3]), El = -w_el R eta (no factor 2), Dis = -w_dis K eta;
   *   precipitate (src/dynamics_precipitate.cu:118-289): eta^2 elastic source (B built from
   *     nvar eigenstrains), F_v += xi 2 R_v eta_v, no elasticity and no stabilisation in the
   *     c update, noise amplitude `noise` on both c and eta (the reference's 1e-4 weightnoise). */
  int rule;
  int chem_form;         /* martensite reference chemical force: 0 HEAD side, 1 master side */
  double w_grad, w_chem, w_dis; /* martensite reference weights                           */
  double temperature, T0;       /* martensite reference: a1 = 0.02 (temperature - T0)       */
  /* Box grids: the reference's `gridsize nx ny nz dx dy dz` (src/dynamics_mart.cu:18) with
   * nx = n and dx as above. 0 means "as n" (ny, nz) or "as dx" (dy, dz). A grid whose edges or
   * spacings are not all equal runs the box path: one kernel per operation on full complex
   * spectra, each axis transformed at its own length (any of the supported edges), FP32 on
   * GSK_PATH_FUSED32 (both rules) and FP64 on GSK_PATH_PLAIN64 (spectral rules); wave numbers
   * per axis (src/gtensorb.cu:14-24), difference quotients per axis (src/dynamics_mart.cu:181-188).
   * Ensembles, the adjoint, the stress transform and per-kernel timing take cubes only. */
  int ny, nz;
  double dy, dz;
} gsk_config;

#define GSK_RULE_SPECTRAL 0
#define GSK_RULE_REFERENCE 1

/* Fill cfg with the defaults of the given physics. */
int gsk_config_default(gsk_config* cfg, int physics);
/* Fill cfg with the defaults of the reference rule of the given physics (rule =
 * GSK_RULE_REFERENCE, kconv = GSK_K_REFERENCE, the plugins' Initialize defaults:
 * src/dynamics_mart.cu:20-41 HEAD side, src/dynamics_precipitate.cu:21-29). */
int gsk_config_reference(gsk_config* cfg, int physics);

int gsk_create(const gsk_config* cfg, gsk_sim** out);
int gsk_destroy(gsk_sim* sim);

/* Khachaturyan B tensor from elastic constants C[3][3][3][3] (row-major, 81
 * doubles) and eigenstrains eps[neps][3][3]. Martensite: neps = nvar, B_pq acts
 * on eta^2. Precipitate: neps = 1, the eigenstrain per unit concentration. */
int gsk_build_btensor(gsk_sim* sim, const double* C81, const double* eps, int neps);

/* Fields: "eta" (nvar*n^3) and "concentration" (n^3), layout [v][x][y][z]. Martensite
 * reference rule, set only: "defect" (n^3, src/dynamics_mart.cu:333-346) and
 * "dislocationstress" (6*n^3 components 00,01,02,11,12,22, src/dynamics_mart.cu:125-139;
 * needs the B tensor built first: K_p = 25 sum_ij sigma_ij eps^p_ij uses its eigenstrains). */
int gsk_set_field(gsk_sim* sim, const char* name, const void* host, int dtype);
int gsk_get_field(gsk_sim* sim, const char* name, void* host, int dtype);
/* Philox initial field lo + (hi-lo)*u, u = (x>>8)*2^-24, counter (voxel/4, field, 0, 1). */
int gsk_init_uniform(gsk_sim* sim, const char* name, double lo, double hi, unsigned long long seed);

int gsk_step(gsk_sim* sim, int nsteps);
int gsk_set_step(gsk_sim* sim, long long step);
long long gsk_get_step(const gsk_sim* sim);
int gsk_set_param(gsk_sim* sim, const char* key, double value);

int gsk_reduce(gsk_sim* sim, int what, double* out);

/* Timing on the handle's stream with HIP events: ms over nsteps steps. */
int gsk_time_steps(gsk_sim* sim, int nsteps, double* ms);
/* Per-kernel timing of nsteps steps: fills up to maxk entries of ms (total ms
 * per kernel slot) and names (comma separated into a caller buffer). */
int gsk_kernel_times(gsk_sim* sim, int nsteps, double* ms, int maxk, int* nk, char* names, int names_len);
/* Algorithmic HBM bytes per step of each kernel slot (same order as gsk_kernel_times). */
int gsk_kernel_bytes(gsk_sim* sim, double* bytes, int maxk, int* nk);

/* Philox4x32-10 block on the host (the generator every noise and initial field uses):
 * out = philox(ctr, key = {seed low word, seed high word}). For documented parameter draws. */
int gsk_philox4x32(const unsigned int ctr[4], unsigned long long seed, unsigned int out[4]);
/* Philox normal field (device generated, copied to host), for bit-exact checks. */
int gsk_noise_field(unsigned long long seed, unsigned int step, unsigned int field, long long nvox, float* host_out,
                    int device);
/* FP64 B tensor evaluated on the device in the full [p][q][x][y][z] layout. */
int gsk_btensor_full(int n, double dx, int kconv, const double* C81, const double* eps1, int V1,
                     const double* eps2, int V2, double* host_out, int device);
/* 3D R2C / C2R of `count` real n^3 fields with the library's FP32 transforms.
 * Spectrum layout [f][kx][ky][kz], kz in [0, n/2] (interleaved re, im floats). */
int gsk_fft_r2c(int n, int count, const float* host_in, float* host_out, int device);
int gsk_fft_c2r(int n, int count, const float* host_in, float* host_out, int device);

/* Same transforms timed against rocFFT (out-of-place, single precision, batch = count).
 * out[0] library R2C ms, out[1] rocFFT R2C ms, out[2] library C2R ms, out[3] rocFFT C2R ms,
 * out[4] max |library - rocFFT| / max |rocFFT| over the R2C spectra. */
int gsk_fft_compare(int n, int count, int reps, double* out, int device);

/* Elastic stress of the order parameters, the stress plugin (src/dynamics_stress.cu:19-194):
 * with the six unit base strains b (single entries 00, 01, 02, 11, 12, 22) and the nvar
 * variant eigenstrains eps, stress_b = -xi IFFT[ sum_v B_bv(k) FFT(s_v) ] / n^3 with
 * s_v = eta_v^2 (source 2, HEAD side) or eta_v (source 1, master side); stress is 6*n^3
 * floats in the component order the martensite plugin reads as "dislocationstress";
 * defect (n^3, may be null) is 1 where any |eta_v| > 0.9. Host buffers, synchronous. */
int gsk_stress(int n, double dx, int kconv, const double* C81, const double* eps, int nvar, double xi, int source,
               const float* eta, float* stress, float* defect, int device);

/* ---- adjoint of the FP32 fused martensite step (nvar <= 4). No counterpart in the
 * reference (it has no reverse mode); it differentiates the step gsk_step runs.
 * gsk_adjoint_forward advances the handle by nsteps from its current state like gsk_step
 * and records a checkpoint (real fields, k-space state, step) every K steps plus the final
 * state; K = 1 stores every state. gsk_adjoint_backward takes the upstream gradient
 * dLoss/deta(T) (nvar*n^3 floats, layout of gsk_get_field) and returns dLoss/deta(0)
 * (grad_eta0, may be null) and grad_params = {dLoss/ds_1..s_nvar, dLoss/dL, dLoss/dkappa}
 * (s_p scales eigenstrain p, so B_pq ~ s_p s_q; evaluated at the built tensor, s = 1).
 * The handle is left at the final state. Noise is included (same Philox stream). */
int gsk_adjoint_forward(gsk_sim* sim, int nsteps, int K);
int gsk_adjoint_backward(gsk_sim* sim, const float* dL_detaT, float* grad_eta0, double* grad_params);
/* out[0] checkpoint bytes (incl. final state), out[1] recompute-buffer bytes, out[2] other
 * work bytes, out[3] forward ms, out[4] backward ms, out[5] stored states. */
int gsk_adjoint_info(const gsk_sim* sim, double* out);
int gsk_adjoint_release(gsk_sim* sim);

/* ---- ensembles: many independent instances stepped together on one GPU.
 * No counterpart in the reference (GS::Run drives one plugin instance at a time,
 * src/gs.cu:168-212); this is the batched driver of the same step. Members are ordinary
 * handles (borrowed from the ensemble: use every gsk_* call on them except gsk_destroy)
 * and must use the FP32 fused path. Schedules: */
#define GSK_ENS_SEQUENTIAL 0 /* each member runs its steps alone, one after another */
#define GSK_ENS_STREAMS 1    /* members spread over `branches` streams, launched eagerly */
#define GSK_ENS_GRAPH 2      /* the branches captured as one hipGraph of steps_per_graph steps */
#define GSK_ENS_BATCHED 3    /* members with one grid size (edge <= 96) and variant count (spectral
                                martensite, 1..4 variants) share every kernel launch (one launch per
                                kernel and size class, grid.z = member); size classes run on parallel
                                streams, the other members as in GSK_ENS_STREAMS over `branches`
                                streams */
#define GSK_ENS_BATCHED_GRAPH 4 /* the batched launches of steps_per_graph steps captured as one
                                   hipGraph and replayed */
typedef struct gsk_ensemble gsk_ensemble;
int gsk_ensemble_create(const gsk_config* cfgs, int count, gsk_ensemble** out);
int gsk_ensemble_destroy(gsk_ensemble* ens);
int gsk_ensemble_size(const gsk_ensemble* ens);
int gsk_ensemble_member(gsk_ensemble* ens, int i, gsk_sim** out);
/* One B tensor per distinct (n, nvar, physics, dx, kconv), shared by the members. */
int gsk_ensemble_build_btensor(gsk_ensemble* ens, const double* C81, const double* eps, int neps);
/* branches = 0: min(members, 8). Defaults: GSK_ENS_GRAPH, 0, 10. */
int gsk_ensemble_set_schedule(gsk_ensemble* ens, int mode, int branches, int steps_per_graph);
int gsk_ensemble_get_schedule(const gsk_ensemble* ens, int* mode, int* branches, int* steps_per_graph);
/* Advance every member by nsteps (synchronous). Results are identical to stepping the
 * members one by one with gsk_step, whatever the schedule. */
int gsk_ensemble_step(gsk_ensemble* ens, int nsteps);
/* Time every candidate schedule over trial_steps steps, keep the fastest, and restore the
 * members' fields, state and step counters. table (optional) receives rows of
 * {mode, branches, steps_per_graph, ms per ensemble step}; *rows their count. */
int gsk_ensemble_autotune(gsk_ensemble* ens, int trial_steps, double* table, int max_rows, int* rows);

/* ---- no-allocation variant of gsk_reduce's real-space reductions. No counterpart in the
 * reference; it serves callers that cannot allocate device memory (tight serving loops, harnesses
 * that forbid hipMalloc): it calls no hipMalloc/hipFree and keeps its partial sums in a fixed
 * 40 KiB device array of the library, one per GPU, that every call on that GPU shares (calls that
 * overlap on one GPU race; run them from one thread or serialise them).
 * d_eta: nvar order-parameter fields of nvox values each, in device memory of `device`, layout
 * [v][voxel] as gsk_get_field's (any nvox > 0, nvar 1..8). what = GSK_RED_MEAN_ETA: out[v] = mean of
 * eta_v; GSK_RED_SUM_ETA2: out[0] = sum over voxels and variants of eta^2 -- the values gsk_reduce
 * returns for a handle holding the same fields, bit for bit. Synchronous on the null stream:
 * the caller's writes to d_eta must be complete. Larger nvar is rejected with GSK_ERR_ARG. */
int gsk_reduce_eta_noalloc(const float* d_eta, int nvar, long long nvox, int what, double* out, int device);

const char* gsk_last_error(void);
const char* gsk_version(void);

#ifdef __cplusplus
}
#endif
#endif /* GSK_H */
