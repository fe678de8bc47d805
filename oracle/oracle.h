/*
 * FP64 CPU oracle for the phase-field time step — TEST INFRASTRUCTURE ONLY.
 *
 * This library restates, in plain C++17/OpenMP and double precision, the
 * algorithms of General-Simulator's phase-field path so that the HIP product
 * can be checked against it. Only tests, the smoke check and the CPU-baseline
 * leg of the benchmark may load it; the product never links or calls it.
 *
 * Reference rules restated (paths relative to the General-Simulator tree):
 *   - wave vectors / unit vectors        src/gtensorb.cu:9-34
 *   - Khachaturyan B_pq(n)               src/gtensorb.cu:55-182
 *   - martensite force terms             src/dynamics_mart.cu:270-303 (HEAD sum over all q)
 *   - precipitate local terms            src/dynamics_precipitate.cu:118-168
 *   - precipitate CH / AC updates        src/dynamics_precipitate.cu:197-275
 *   - field indexing [v][x][y][z]        include/datamain.th:70-107
 *   - FP32 total (here FP64)             include/datamain.th:782-806
 * The semi-implicit Allen-Cahn rule used for the martensite plugin follows the
 * precipitate AC form (src/dynamics_precipitate.cu:244-247) with the variational
 * factor 2 of src/dynamics_precipitate.cu:193.
 *
 * cuFFT / cuRAND (closed NVIDIA libraries linked by the reference) are replaced:
 * the FFT here is an independent self-sorting Stockham (radices 4, 2, 3) in
 * double precision; noise is Philox4x32-10 (Salmon et al., SC'11) followed by
 * the documented fp32 Box-Muller transform built from + - * only.
 */
#ifndef GS_ORACLE_H
#define GS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_mart_params {
  double A, B, C;        /* Landau: f = A/2 sum eta^2 - B/3 sum eta^3 + C/4 (sum eta^2)^2 */
  double kappa, L, dt;   /* gradient coefficient, mobility, time step                    */
  double w_el;           /* elastic weight                                               */
  double noise;          /* noise amplitude a (force term a*xi)                          */
  double dx;             /* grid spacing                                                 */
  int kconv;             /* 0: k = 2 pi m/(n dx); 1: reference form 2*3.14*dx*m/n         */
  uint64_t seed;
} or_mart_params;

typedef struct or_prec_params {
  double A[8];           /* A[1..7] as in the reference coefficient list                 */
  double c1, c2;         /* concentrations                                               */
  double M, L;           /* CH and AC mobilities                                         */
  double kappa_c, kappa_eta; /* beta and alpha (gradient coefficients)                   */
  double xi;             /* elastic weight                                               */
  double S;              /* CH stabilisation: S*k^2 moved to the implicit side (0 = reference rule) */
  double dt, noise, dx;  /* time step, eta-noise amplitude, grid spacing                 */
  int kconv;
  uint64_t seed;
} or_prec_params;

/* Reference rule of the martensite plugin (explicit Euler, src/dynamics_mart.cu:181-417). */
typedef struct or_mart_ref_params {
  double a1, a2, a3;     /* LPC[1..3]; a1 = 0.02 (T - T0) at the step                    */
  double w_grad, w_chem, w_el, w_dis, w_noise;
  double dt, dx;
  int kconv;
  int chem_form;         /* 0: HEAD side of ChemiFor_Mart_Kernel, 1: master side         */
  uint64_t seed;
} or_mart_ref_params;
/* defect (n^3) and K (dislocation force constants, V*n^3) may be null. */
void or_mart_ref_run(int n, int V, const or_mart_ref_params* p, const double* B, double* eta,
                     const double* defect, const double* K, int step0, int nsteps);
/* Reference rule of the precipitate plugin: eta^2 elastic source with the full V x V B. */
void or_prec_ref_run(int n, int V, const or_prec_params* p, const double* B, double* c, double* eta,
                     int step0, int nsteps);

/* Stress plugin: B is the (6 base strains) x V tensor [6][V][n^3]; stress 6*n^3, defect n^3. */
void or_stress(int n, int V, const double* B, double xi, int source, const double* eta, double* stress,
               double* defect);

/* Box grids: nx x ny x nz voxels with spacings dx, dy, dz (the reference's `gridsize nx ny nz dx dy
 * dz`, src/dynamics_mart.cu:18; per-axis wave numbers src/gtensorb.cu:14-24, per-axis difference
 * quotients src/dynamics_mart.cu:181-188). The _box entry points take the grid in place of n and
 * use its spacings in place of the params' dx. */
typedef struct or_grid {
  int nx, ny, nz;
  double dx, dy, dz;
} or_grid;
void or_fft3d_box(const or_grid* g, double* data, int sign);
void or_btensor_box(const or_grid* g, int kconv, const double C[81], const double* eps1, int V1,
                    const double* eps2, int V2, double* B);
void or_mart_run_box(const or_grid* g, int V, const or_mart_params* p, const double* B, double* eta, int step0,
                     int nsteps);
void or_mart_energy_box(const or_grid* g, int V, const or_mart_params* p, const double* B, const double* eta,
                        double out[4]);
void or_mart_ref_run_box(const or_grid* g, int V, const or_mart_ref_params* p, const double* B, double* eta,
                         const double* defect, const double* K, int step0, int nsteps);
void or_prec_run_box(const or_grid* g, int V, const or_prec_params* p, const double* Bc, double* c, double* eta,
                     int step0, int nsteps);
void or_prec_energy_box(const or_grid* g, int V, const or_prec_params* p, const double* Bc, const double* c,
                        const double* eta, double out[4]);
void or_prec_ref_run_box(const or_grid* g, int V, const or_prec_params* p, const double* B, double* c,
                         double* eta, int step0, int nsteps);

void or_set_threads(int nt);
int or_get_threads(void);

/* Complex FFTs on interleaved (re,im) doubles; sign=-1 forward, +1 backward; unnormalised. */
void or_fft1d(int n, double* data, int sign);
void or_fft3d(int n, double* data, int sign);

/* Philox4x32-10 block function and the normal transform. */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void or_normal4(const uint32_t x[4], float out[4]);
float or_log_f(float u);
float or_sqrt_f(float x);
void or_sincos_f(float theta, float* s, float* c);
void or_noise_field(uint64_t seed, uint32_t step, uint32_t field, int64_t nvox, float* out);
void or_uniform_field(uint64_t seed, uint32_t field, int64_t nvox, double lo, double hi, float* out);

/* Wave vector squared |k|^2 for index m of an n-point axis. */
double or_kcomp(int m, int n, double dx, int kconv);

/* B[p][q][x][y][z] (V1*V2*n^3 doubles). eps1: V1*9, eps2: V2*9 (row-major 3x3). */
void or_btensor(int n, double dx, int kconv, const double C[81],
                const double* eps1, int V1, const double* eps2, int V2, double* B);

/* Martensite semi-implicit AC. eta: V*n^3 doubles, advanced in place from step0. */
void or_mart_run(int n, int V, const or_mart_params* p, const double* B,
                 double* eta, int step0, int nsteps);
/* out[0]=chemical, out[1]=gradient, out[2]=elastic, out[3]=total */
void or_mart_energy(int n, int V, const or_mart_params* p, const double* B,
                    const double* eta, double out[4]);
/* Store-all adjoint of or_mart_run over nsteps from eta0 (step index step0) for the upstream
 * gradient gT = dLoss/deta(T): g_eta0 = dLoss/deta0 (V*n^3), g_par = {dLoss/ds_1..s_V
 * (eigenstrain magnitudes, B_pq ~ s_p s_q, at s = 1), dLoss/dL, dLoss/dkappa}. */
void or_mart_adjoint(int n, int V, const or_mart_params* p, const double* B, const double* eta0, int step0,
                     int nsteps, const double* gT, double* g_eta0, double* g_par);
/* Force F_p of the martensite rule (no noise) for checking against energy derivatives. */
void or_mart_force(int n, int V, const or_mart_params* p, const double* B,
                   const double* eta, double* F);

/* Precipitate CH + AC with the dilatational c-sourced elasticity (scalar Bc[n^3]). */
void or_prec_run(int n, int V, const or_prec_params* p, const double* Bc,
                 double* c, double* eta, int step0, int nsteps);
void or_prec_energy(int n, int V, const or_prec_params* p, const double* Bc,
                    const double* c, const double* eta, double out[4]);
void or_prec_local(int V, const or_prec_params* p, double c, const double* eta,
                   double* mu_c, double* F, double* f);

/* FP64 sums */
double or_sum(const double* a, int64_t n);

#ifdef __cplusplus
}
#endif
#endif
