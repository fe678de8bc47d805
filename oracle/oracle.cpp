// FP64 CPU oracle — TEST INFRASTRUCTURE ONLY (see oracle.h for scope and citations).
// Built with -ffp-contract=off so that the fp32 normal transform rounds exactly
// as written, operation by operation.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef std::complex<double> cd;
static const double kPi = 3.14159265358979323846;

static int g_threads = 0;
extern "C" void or_set_threads(int nt) {
  g_threads = nt;
#ifdef _OPENMP
  if (nt > 0) omp_set_num_threads(nt);
#endif
}
extern "C" int or_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// ---------------------------------------------------------------------------
// FFT: self-sorting Stockham passes with radices 4, 2 and 3, on batches of B lines
// stored line-fastest ([position][line]) so that the strided y and x passes of a 3D
// transform read and write runs of B contiguous z values. One pass of radix R with
// Ns = product of the radices before it:
//   for j < n/R:  k = j mod Ns,  v_r = in[j + r n/R] * exp(sign 2 pi i k r / (Ns R)),
//                 X_q = sum_r v_r exp(sign 2 pi i r q / R),  out[(j/Ns) Ns R + k + q Ns] = X_q.
// Twiddles are evaluated with cos/sin in double precision for every (pass, k, r).
// ---------------------------------------------------------------------------
struct FftPlan {
  int n = 0, sign = -1;
  std::vector<int> radix;
  std::vector<std::vector<cd>> tw;  // tw[pass][k * R + r]
  FftPlan(int n_, int sign_) : n(n_), sign(sign_) {
    int m = n;
    while (m % 4 == 0) { radix.push_back(4); m /= 4; }
    while (m % 2 == 0) { radix.push_back(2); m /= 2; }
    while (m % 3 == 0) { radix.push_back(3); m /= 3; }
    if (m != 1) radix.push_back(m);  // any other factor: one direct DFT pass
    int Ns = 1;
    for (int R : radix) {
      std::vector<cd> t((size_t)Ns * R);
      for (int k = 0; k < Ns; ++k)
        for (int r = 0; r < R; ++r) {
          const double a = sign * 2.0 * kPi * (double)k * (double)r / (double)(Ns * R);
          t[(size_t)k * R + r] = cd(std::cos(a), std::sin(a));
        }
      tw.push_back(std::move(t));
      Ns *= R;
    }
  }
  // data, scratch: n * B complex values, element (i, b) at i * B + b. Result in data.
  void run(cd* data, cd* scratch, int B) const {
    cd* in = data;
    cd* out = scratch;
    int Ns = 1;
    const double s = (double)sign;
    const double c3 = -0.5, s3 = s * 0.86602540378443864676;  // exp(sign 2 pi i / 3)
    for (size_t ps = 0; ps < radix.size(); ++ps) {
      const int R = radix[ps], nR = n / R;
      const cd* w = tw[ps].data();
      for (int j = 0; j < nR; ++j) {
        const int k = j % Ns;
        const int dst = (j / Ns) * Ns * R + k;
        const cd* wk = w + (size_t)k * R;
        for (int b = 0; b < B; ++b) {
          double vr[16], vi[16];
          for (int r = 0; r < R; ++r) {
            const cd x = in[(size_t)(j + r * nR) * B + b];
            vr[r] = x.real() * wk[r].real() - x.imag() * wk[r].imag();
            vi[r] = x.real() * wk[r].imag() + x.imag() * wk[r].real();
          }
          double xr[16], xi[16];
          if (R == 4) {
            const double ar = vr[0] + vr[2], ai = vi[0] + vi[2], br = vr[0] - vr[2], bi = vi[0] - vi[2];
            const double cr = vr[1] + vr[3], ci = vi[1] + vi[3], dr = vr[1] - vr[3], di = vi[1] - vi[3];
            // exp(sign 2 pi i / 4) = sign * i:  (sign i) d = (-sign di, sign dr)
            xr[0] = ar + cr; xi[0] = ai + ci;
            xr[2] = ar - cr; xi[2] = ai - ci;
            xr[1] = br - s * di; xi[1] = bi + s * dr;
            xr[3] = br + s * di; xi[3] = bi - s * dr;
          } else if (R == 2) {
            xr[0] = vr[0] + vr[1]; xi[0] = vi[0] + vi[1];
            xr[1] = vr[0] - vr[1]; xi[1] = vi[0] - vi[1];
          } else if (R == 3) {
            const double tr = vr[1] + vr[2], ti = vi[1] + vi[2], ur = vr[1] - vr[2], ui = vi[1] - vi[2];
            xr[0] = vr[0] + tr; xi[0] = vi[0] + ti;
            const double mr = vr[0] + c3 * tr, mi = vi[0] + c3 * ti;
            xr[1] = mr - s3 * ui; xi[1] = mi + s3 * ur;
            xr[2] = mr + s3 * ui; xi[2] = mi - s3 * ur;
          } else {
            for (int q = 0; q < R; ++q) {
              double accr = 0.0, acci = 0.0;
              for (int r = 0; r < R; ++r) {
                const double a = s * 2.0 * kPi * (double)((r * q) % R) / (double)R;
                const double cr = std::cos(a), ci = std::sin(a);
                accr += vr[r] * cr - vi[r] * ci;
                acci += vr[r] * ci + vi[r] * cr;
              }
              xr[q] = accr; xi[q] = acci;
            }
          }
          for (int q = 0; q < R; ++q) out[(size_t)(dst + q * Ns) * B + b] = cd(xr[q], xi[q]);
        }
      }
      std::swap(in, out);
      Ns *= R;
    }
    if (in != data) std::memcpy(data, in, sizeof(cd) * (size_t)n * B);
  }
};

extern "C" void or_fft1d(int n, double* data, int sign) {
  FftPlan pl(n, sign);
  std::vector<cd> s(n);
  pl.run(reinterpret_cast<cd*>(data), s.data(), 1);
}

// A periodic grid of nx x ny x nz voxels with spacings dx, dy, dz, fields stored [x][y][z] (z
// fastest, include/datamain.th:70-76); the reference reads it from `gridsize nx ny nz dx dy dz`
// (src/dynamics_mart.cu:18). The cube entry points use nx = ny = nz = n and one spacing.
struct Grid {
  int nx, ny, nz;
  double dx, dy, dz;
  int64_t N;
  void coords(int64_t i, int& x, int& y, int& z) const {
    x = (int)(i / ((int64_t)ny * nz));
    y = (int)((i / nz) % ny);
    z = (int)(i % nz);
  }
};
static Grid cube_grid(int n, double dx) { return Grid{n, n, n, dx, dx, dx, (int64_t)n * n * n}; }
static Grid box_grid(const or_grid* g) {
  return Grid{g->nx, g->ny, g->nz, g->dx, g->dy, g->dz, (int64_t)g->nx * g->ny * g->nz};
}

// 3D transform as 1D transforms along z, y and x, each on batches of up to 16 lines (the result
// does not depend on the batching: every line is transformed on its own).
static void fft3d(const Grid& g, cd* a, int sign) {
  const FftPlan px(g.nx, sign), py(g.ny, sign), pz(g.nz, sign);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int64_t sx = (int64_t)ny * nz, sy = nz;
  const int B = 16;
  const int nby = (ny + B - 1) / B, nbz = (nz + B - 1) / B;
  const int L = std::max(nx, std::max(ny, nz));
#pragma omp parallel
  {
    std::vector<cd> buf((size_t)L * B), s((size_t)L * B);
    // z lines (contiguous): B lines of one (x, y-block) gathered transposed
#pragma omp for schedule(static)
    for (int64_t t = 0; t < (int64_t)nx * nby; ++t) {
      const int64_t x = t / nby, y0 = (t % nby) * B;
      const int Bt = (int)std::min<int64_t>(B, ny - y0);
      for (int b = 0; b < Bt; ++b)
        for (int z = 0; z < nz; ++z) buf[(size_t)z * Bt + b] = a[x * sx + (y0 + b) * sy + z];
      pz.run(buf.data(), s.data(), Bt);
      for (int b = 0; b < Bt; ++b)
        for (int z = 0; z < nz; ++z) a[x * sx + (y0 + b) * sy + z] = buf[(size_t)z * Bt + b];
    }
    // y lines: for each x and block of B z values, element (y, b) = a[x][y][z0 + b]
#pragma omp for schedule(static)
    for (int64_t t = 0; t < (int64_t)nx * nbz; ++t) {
      const int64_t x = t / nbz, z0 = (t % nbz) * B;
      const int Bt = (int)std::min<int64_t>(B, nz - z0);
      for (int y = 0; y < ny; ++y)
        for (int b = 0; b < Bt; ++b) buf[(size_t)y * Bt + b] = a[x * sx + (int64_t)y * sy + z0 + b];
      py.run(buf.data(), s.data(), Bt);
      for (int y = 0; y < ny; ++y)
        for (int b = 0; b < Bt; ++b) a[x * sx + (int64_t)y * sy + z0 + b] = buf[(size_t)y * Bt + b];
    }
    // x lines: for each y and block of B z values, element (x, b) = a[x][y][z0 + b]
#pragma omp for schedule(static)
    for (int64_t t = 0; t < (int64_t)ny * nbz; ++t) {
      const int64_t y = t / nbz, z0 = (t % nbz) * B;
      const int Bt = (int)std::min<int64_t>(B, nz - z0);
      for (int x = 0; x < nx; ++x)
        for (int b = 0; b < Bt; ++b) buf[(size_t)x * Bt + b] = a[(int64_t)x * sx + y * sy + z0 + b];
      px.run(buf.data(), s.data(), Bt);
      for (int x = 0; x < nx; ++x)
        for (int b = 0; b < Bt; ++b) a[(int64_t)x * sx + y * sy + z0 + b] = buf[(size_t)x * Bt + b];
    }
  }
}

extern "C" void or_fft3d(int n, double* data, int sign) { fft3d(cube_grid(n, 1.0), reinterpret_cast<cd*>(data), sign); }
extern "C" void or_fft3d_box(const or_grid* g, double* data, int sign) {
  fft3d(box_grid(g), reinterpret_cast<cd*>(data), sign);
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): 10 rounds of
//   (hi0,lo0)=M0*c0, (hi1,lo1)=M1*c2, c=(hi1^c1^k0, lo1, hi0^c3^k1, lo0),
// key bumped by the Weyl constants between rounds.
// ---------------------------------------------------------------------------
extern "C" void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// ---------------------------------------------------------------------------
// fp32 Box-Muller built from + - * and integer bit operations only, so that any
// IEEE-754 binary32 implementation with round-to-nearest-even and no operation
// contraction produces the same bits. Constants are given as bit patterns.
// ---------------------------------------------------------------------------
static inline float f_of(uint32_t b) { float f; std::memcpy(&f, &b, 4); return f; }
static inline uint32_t b_of(float f) { uint32_t b; std::memcpy(&b, &f, 4); return b; }

extern "C" float or_log_f(float u) {
  // u in [2^-24, 1]; ln(u) = e*ln2 + ln(1+f), m = 1+f in [sqrt(1/2), sqrt(2)]
  uint32_t i = b_of(u);
  int e = (int)((i >> 23) & 0xFFu) - 127;
  float m = f_of((i & 0x007FFFFFu) | 0x3F800000u);
  if (m > f_of(0x3FB504F3u)) { m = m * 0.5f; e = e + 1; }
  float f = m - 1.0f;
  // ln(1+f) = f * q(f), q of degree 9 (least-squares fit on [sqrt(1/2)-1, sqrt(2)-1])
  float q = f_of(0xBD9C70BBu);
  q = q * f + f_of(0x3E043E25u);
  q = q * f + f_of(0xBE07962Fu);
  q = q * f + f_of(0x3E11342Cu);
  q = q * f + f_of(0xBE2A13CFu);
  q = q * f + f_of(0x3E4CD23Cu);
  q = q * f + f_of(0xBE800214u);
  q = q * f + f_of(0x3EAAAAA8u);
  q = q * f + f_of(0xBEFFFFFCu);
  q = q * f + f_of(0x3F800000u);
  float p = f * q;
  float t = (float)e;
  return t * f_of(0x3F317200u) + (p + t * f_of(0x35BFBE8Eu));
}

extern "C" float or_sqrt_f(float x) {
  if (!(x > 0.0f)) return 0.0f;
  uint32_t i = b_of(x);
  float y = f_of(0x5F3759DFu - (i >> 1));
  float h = 0.5f * x;
  for (int it = 0; it < 3; ++it) {
    float t = y * y;
    t = h * t;
    t = 1.5f - t;
    y = y * t;
  }
  return x * y;
}

extern "C" void or_sincos_f(float th, float* s, float* c) {
  // theta in [0, pi/2): Taylor series to theta^13 (sin) and theta^14 (cos)
  float z = th * th;
  float ps = f_of(0x2F309231u);
  ps = ps * z + f_of(0xB2D7322Bu);
  ps = ps * z + f_of(0x3638EF1Du);
  ps = ps * z + f_of(0xB9500D01u);
  ps = ps * z + f_of(0x3C088889u);
  ps = ps * z + f_of(0xBE2AAAABu);
  ps = ps * z + 1.0f;
  *s = th * ps;
  float pc = f_of(0xAD49CBA5u);
  pc = pc * z + f_of(0x310F76C7u);
  pc = pc * z + f_of(0xB493F27Eu);
  pc = pc * z + f_of(0x37D00D01u);
  pc = pc * z + f_of(0xBAB60B61u);
  pc = pc * z + f_of(0x3D2AAAABu);
  pc = pc * z + f_of(0xBF000000u);
  *c = pc * z + 1.0f;
}

static void normal_pair(uint32_t a, uint32_t b, float* z0, float* z1) {
  float u1 = (float)((a >> 8) + 1u) * f_of(0x33800000u);  // * 2^-24, in (0,1]
  float L = or_log_f(u1);
  float r = or_sqrt_f(-2.0f * L);
  uint32_t q = b >> 30;
  float fr = (float)((b >> 8) & 0x3FFFFFu) * f_of(0x34800000u);  // * 2^-22, in [0,1)
  float th = fr * f_of(0x3FC90FDBu);                                // * pi/2
  float s, c;
  or_sincos_f(th, &s, &c);
  float cc, ss;
  switch (q) {
    case 0: cc = c; ss = s; break;
    case 1: cc = -s; ss = c; break;
    case 2: cc = -c; ss = -s; break;
    default: cc = s; ss = -c; break;
  }
  *z0 = r * cc;
  *z1 = r * ss;
}

extern "C" void or_normal4(const uint32_t x[4], float out[4]) {
  normal_pair(x[0], x[1], &out[0], &out[1]);
  normal_pair(x[2], x[3], &out[2], &out[3]);
}

// counter = (voxel/4, field, step, purpose); key = (seed low, seed high)
extern "C" void or_noise_field(uint64_t seed, uint32_t step, uint32_t field, int64_t nvox, float* out) {
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
#pragma omp parallel for schedule(static)
  for (int64_t g = 0; g < nvox / 4; ++g) {
    uint32_t ctr[4] = {(uint32_t)g, field, step, 0u}, x[4];
    or_philox4x32_10(ctr, key, x);
    or_normal4(x, out + 4 * g);
  }
}

extern "C" void or_uniform_field(uint64_t seed, uint32_t field, int64_t nvox, double lo, double hi, float* out) {
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
#pragma omp parallel for schedule(static)
  for (int64_t g = 0; g < nvox / 4; ++g) {
    uint32_t ctr[4] = {(uint32_t)g, field, 0u, 1u}, x[4];
    or_philox4x32_10(ctr, key, x);
    for (int i = 0; i < 4; ++i) {
      double u = (double)(x[i] >> 8) * (1.0 / 16777216.0);
      out[4 * g + i] = (float)(lo + (hi - lo) * u);
    }
  }
}

// ---------------------------------------------------------------------------
// Wave vectors and the Khachaturyan tensor (src/gtensorb.cu:9-34, 110-182)
// ---------------------------------------------------------------------------
extern "C" double or_kcomp(int m, int n, double dx, int kconv) {
  int t = (m < n / 2) ? m : m - n;  // src/gtensorb.cu:10
  if (kconv == 1) return 2.0 * 3.14 * dx / n * t;  // src/gtensorb.cu:11
  return 2.0 * kPi * t / (n * dx);
}

static void inv3(const double a[9], double o[9]) {
  double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
               a[2] * (a[3] * a[7] - a[4] * a[6]);
  if (std::fabs(det) < 1e-8) {  // src/gtensorb.cu:145,155-156
    for (int i = 0; i < 9; ++i) o[i] = 0.0;
    return;
  }
  o[0] = (a[4] * a[8] - a[5] * a[7]) / det;
  o[1] = (a[2] * a[7] - a[1] * a[8]) / det;
  o[2] = (a[1] * a[5] - a[2] * a[4]) / det;
  o[3] = (a[5] * a[6] - a[3] * a[8]) / det;
  o[4] = (a[0] * a[8] - a[2] * a[6]) / det;
  o[5] = (a[2] * a[3] - a[0] * a[5]) / det;
  o[6] = (a[3] * a[7] - a[4] * a[6]) / det;
  o[7] = (a[1] * a[6] - a[0] * a[7]) / det;
  o[8] = (a[0] * a[4] - a[1] * a[3]) / det;
}

// B over a grid: the per-axis wave numbers of GTensorB_VF with each axis's own n and spacing
// (src/gtensorb.cu:14-24), the unit vector and the tensor of src/gtensorb.cu:110-182.
static void btensor(const Grid& g, int kconv, const double C[81], const double* eps1, int V1, const double* eps2,
                    int V2, double* B) {
  const int64_t N = g.N;
  // sigma = C : eps  (src/gtensorb.cu:72-84)
  std::vector<double> s1(V1 * 9, 0.0), s2(V2 * 9, 0.0), cee(V1 * V2, 0.0);
  for (int p = 0; p < V1; ++p)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l) s1[p * 9 + i * 3 + j] += C[((i * 3 + j) * 3 + k) * 3 + l] * eps1[p * 9 + k * 3 + l];
  for (int q = 0; q < V2; ++q)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l) s2[q * 9 + i * 3 + j] += C[((i * 3 + j) * 3 + k) * 3 + l] * eps2[q * 9 + k * 3 + l];
  for (int p = 0; p < V1; ++p)
    for (int q = 0; q < V2; ++q) {
      double acc = 0;
      for (int ij = 0; ij < 9; ++ij)
        for (int kl = 0; kl < 9; ++kl) acc += C[ij * 9 + kl] * eps1[p * 9 + ij] * eps2[q * 9 + kl];
      cee[p * V2 + q] = acc;
    }
#pragma omp parallel for schedule(static)
  for (int64_t idx = 0; idx < N; ++idx) {
    int x, y, z;
    g.coords(idx, x, y, z);
    double k[3] = {or_kcomp(x, g.nx, g.dx, kconv), or_kcomp(y, g.ny, g.dy, kconv), or_kcomp(z, g.nz, g.dz, kconv)};
    double mo = std::sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
    double u[3] = {0, 0, 0};
    if (mo > 5e-7) { u[0] = k[0] / mo; u[1] = k[1] / mo; u[2] = k[2] / mo; }  // src/gtensorb.cu:25-33
    double om_inv[9], om[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double acc = 0;
        for (int kk = 0; kk < 3; ++kk)
          for (int l = 0; l < 3; ++l) acc += C[((i * 3 + kk) * 3 + l) * 3 + j] * u[kk] * u[l];  // :127-130
        om_inv[i * 3 + j] = acc;
      }
    inv3(om_inv, om);
    for (int p = 0; p < V1; ++p) {
      double a[3];  // a_j = n_i sigma^p_ij
      for (int j = 0; j < 3; ++j) a[j] = u[0] * s1[p * 9 + j] + u[1] * s1[p * 9 + 3 + j] + u[2] * s1[p * 9 + 6 + j];
      for (int q = 0; q < V2; ++q) {
        double b[3];  // b_k = sigma^q_kl n_l
        for (int kk = 0; kk < 3; ++kk)
          b[kk] = s2[q * 9 + kk * 3] * u[0] + s2[q * 9 + kk * 3 + 1] * u[1] + s2[q * 9 + kk * 3 + 2] * u[2];
        double t = 0;
        for (int j = 0; j < 3; ++j)
          for (int kk = 0; kk < 3; ++kk) t += a[j] * om[j * 3 + kk] * b[kk];
        B[(int64_t)(p * V2 + q) * N + idx] = cee[p * V2 + q] - t;  // :168-178
      }
    }
  }
}

extern "C" void or_btensor(int n, double dx, int kconv, const double C[81], const double* eps1, int V1,
                           const double* eps2, int V2, double* B) {
  btensor(cube_grid(n, dx), kconv, C, eps1, V1, eps2, V2, B);
}
extern "C" void or_btensor_box(const or_grid* g, int kconv, const double C[81], const double* eps1, int V1,
                               const double* eps2, int V2, double* B) {
  btensor(box_grid(g), kconv, C, eps1, V1, eps2, V2, B);
}

// ---------------------------------------------------------------------------
// Martensite: semi-implicit spectral Allen-Cahn with Khachaturyan elasticity
// ---------------------------------------------------------------------------
static std::vector<double> ksq_table(int n, double dx, int kconv) {
  std::vector<double> k(n);
  for (int m = 0; m < n; ++m) { double v = or_kcomp(m, n, dx, kconv); k[m] = v * v; }
  return k;
}
// |k|^2 of every voxel index of a grid (each axis with its own n and spacing)
struct KSq {
  std::vector<double> x, y, z;
  const Grid* g;
  KSq(const Grid& gr, int kconv)
      : x(ksq_table(gr.nx, gr.dx, kconv)), y(ksq_table(gr.ny, gr.dy, kconv)), z(ksq_table(gr.nz, gr.dz, kconv)),
        g(&gr) {}
  double operator()(int64_t i) const {
    int a, b, c;
    g->coords(i, a, b, c);
    return x[a] + y[b] + z[c];
  }
};

// R_p = IFFT[ sum_q B_pq FFT(eta_q^2) ] / N   (src/dynamics_mart.cu:283-289, HEAD: all q)
static void elastic_R(const Grid& g, int V, const double* B, const double* eta, std::vector<cd>& R) {
  const int64_t N = g.N;
  std::vector<cd> S((size_t)V * N);
  for (int q = 0; q < V; ++q) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) { double e = eta[q * N + i]; S[q * N + i] = cd(e * e, 0.0); }
    fft3d(g, S.data() + q * N, -1);
  }
  R.assign((size_t)V * N, cd(0, 0));
  for (int p = 0; p < V; ++p) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
      cd acc = 0;
      for (int q = 0; q < V; ++q) acc += B[(int64_t)(p * V + q) * N + i] * S[q * N + i];
      R[p * N + i] = acc / (double)N;
    }
    fft3d(g, R.data() + p * N, +1);
  }
}

static void mart_force(const Grid& g, int V, const or_mart_params* p, const double* B, const double* eta, double* F) {
  const int64_t N = g.N;
  std::vector<cd> R;
  elastic_R(g, V, B, eta, R);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    double s2 = 0;
    for (int v = 0; v < V; ++v) s2 += eta[v * N + i] * eta[v * N + i];
    for (int v = 0; v < V; ++v) {
      double e = eta[v * N + i];
      F[v * N + i] = p->A * e - p->B * e * e + p->C * e * s2 + 2.0 * p->w_el * e * R[v * N + i].real();
    }
  }
}
extern "C" void or_mart_force(int n, int V, const or_mart_params* p, const double* B, const double* eta, double* F) {
  mart_force(cube_grid(n, p->dx), V, p, B, eta, F);
}

static void mart_run(const Grid& g, int V, const or_mart_params* p, const double* B, double* eta, int step0,
                     int nsteps) {
  const int64_t N = g.N;
  std::vector<double> F((size_t)V * N);
  const KSq kk(g, p->kconv);
  std::vector<float> xi(N);
  std::vector<cd> eh(N), fh(N);
  for (int s = 0; s < nsteps; ++s) {
    mart_force(g, V, p, B, eta, F.data());
    for (int v = 0; v < V; ++v) {
      if (p->noise != 0.0) {
        or_noise_field(p->seed, (uint32_t)(step0 + s), (uint32_t)v, N, xi.data());
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; ++i) F[v * N + i] += p->noise * (double)xi[i];
      }
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < N; ++i) { eh[i] = cd(eta[v * N + i], 0); fh[i] = cd(F[v * N + i], 0); }
      fft3d(g, eh.data(), -1);
      fft3d(g, fh.data(), -1);
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < N; ++i)
        eh[i] = (eh[i] - p->dt * p->L * fh[i]) / (1.0 + p->dt * p->L * p->kappa * kk(i)) / (double)N;
      fft3d(g, eh.data(), +1);
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < N; ++i) eta[v * N + i] = eh[i].real();
    }
  }
}
extern "C" void or_mart_run(int n, int V, const or_mart_params* p, const double* B, double* eta, int step0,
                            int nsteps) {
  mart_run(cube_grid(n, p->dx), V, p, B, eta, step0, nsteps);
}
extern "C" void or_mart_run_box(const or_grid* g, int V, const or_mart_params* p, const double* B, double* eta,
                                int step0, int nsteps) {
  mart_run(box_grid(g), V, p, B, eta, step0, nsteps);
}

// Store-all adjoint of or_mart_run (FP64 reference gradient for the GPU adjoint; the
// reference itself has no adjoint). One step is eta' = D * (eta - dt L F(eta)) with
// D(k) = 1/(1 + dt L kappa k^2); for the upstream gradient lam = dLoss/deta':
//   mu      = D * lam
//   dLoss/deta_p = mu_p - dt L [ (A - 2B eta_p + C S + 2w R_p) mu_p + 2C eta_p sum_q eta_q mu_q
//                               + 4w eta_p Q_p ],   Q_p = IFFT[sum_r B_pr FFT(eta_r mu_r)]/N
//   dLoss/dL     += <lam, dD/dL D^-1 eta'> - dt <mu, F>,  dD/dL = -dt kappa k^2 D^2
//   dLoss/dkappa += <lam, dD/dkappa D^-1 eta'>,           dD/dkappa = -dt L k^2 D^2
//   dLoss/ds_p   += -2 dt L w sum_q (G_pq + G_qp),  G_rq = <eta_r mu_r, B_rq * eta_q^2>
// where s_p scales eigenstrain p (B_pq proportional to s_p s_q), evaluated at s = 1.
// g_par = {ds_1..ds_V, dL, dkappa}.
extern "C" void or_mart_adjoint(int n, int V, const or_mart_params* p, const double* B, const double* eta0, int step0,
                                int nsteps, const double* gT, double* g_eta0, double* g_par) {
  const Grid g = cube_grid(n, p->dx);
  const int64_t N = g.N, VN = (int64_t)V * N;
  std::vector<double> traj((size_t)(nsteps + 1) * VN);
  std::memcpy(traj.data(), eta0, VN * sizeof(double));
  for (int s = 0; s < nsteps; ++s) {
    std::memcpy(traj.data() + (s + 1) * VN, traj.data() + s * VN, VN * sizeof(double));
    mart_run(g, V, p, B, traj.data() + (s + 1) * VN, step0 + s, 1);
  }
  const std::vector<double> k2 = ksq_table(n, p->dx, p->kconv);
  auto kk_of = [&](int64_t i) {
    const int x = (int)(i / ((int64_t)n * n)), y = (int)((i / n) % n), z = (int)(i % n);
    return k2[x] + k2[y] + k2[z];
  };
  const double dtL = p->dt * p->L, w = p->w_el;
  std::vector<double> lam(gT, gT + VN), mu(VN), F(VN), gs(V, 0.0);
  std::vector<float> xi(N);
  std::vector<cd> a(N), b(N), R, Q((size_t)VN), Sh((size_t)VN), Ph((size_t)VN);
  double gL = 0.0, gk = 0.0;
  for (int s = nsteps - 1; s >= 0; --s) {
    const double* e = traj.data() + s * VN;
    const double* e1 = traj.data() + (s + 1) * VN;
    double X = 0.0;  // (1/N) sum_k Re(conj(lam_hat) eta1_hat) k^2 D
    for (int v = 0; v < V; ++v) {
      for (int64_t i = 0; i < N; ++i) { a[i] = cd(lam[v * N + i], 0); b[i] = cd(e1[v * N + i], 0); }
      fft3d(g, a.data(), -1);
      fft3d(g, b.data(), -1);
      double x = 0.0;
#pragma omp parallel for reduction(+ : x) schedule(static)
      for (int64_t i = 0; i < N; ++i) {
        const double kk = kk_of(i), D = 1.0 / (1.0 + dtL * p->kappa * kk);
        x += (std::conj(a[i]) * b[i]).real() * kk * D;
        a[i] *= D / (double)N;
      }
      X += x / (double)N;
      fft3d(g, a.data(), +1);
      for (int64_t i = 0; i < N; ++i) mu[v * N + i] = a[i].real();
    }
    gL += -p->dt * p->kappa * X;
    gk += -dtL * X;
    elastic_R(g, V, B, e, R);
    // spectra of eta_q^2 and eta_r mu_r
    for (int q = 0; q < V; ++q) {
      for (int64_t i = 0; i < N; ++i) {
        const double ev = e[q * N + i];
        Sh[q * N + i] = cd(ev * ev, 0);
        Ph[q * N + i] = cd(ev * mu[q * N + i], 0);
      }
      fft3d(g, Sh.data() + q * N, -1);
      fft3d(g, Ph.data() + q * N, -1);
    }
    for (int r = 0; r < V; ++r)
      for (int q = 0; q < V; ++q) {
        double G = 0.0;
        const double* Brq = B + (int64_t)(r * V + q) * N;
#pragma omp parallel for reduction(+ : G) schedule(static)
        for (int64_t i = 0; i < N; ++i) G += (std::conj(Ph[r * N + i]) * Brq[i] * Sh[q * N + i]).real();
        G /= (double)N;
        gs[r] += -2.0 * dtL * w * G;
        gs[q] += -2.0 * dtL * w * G;
      }
    for (int pp = 0; pp < V; ++pp) {
      for (int64_t i = 0; i < N; ++i) {
        cd acc = 0;
        for (int r = 0; r < V; ++r) acc += B[(int64_t)(pp * V + r) * N + i] * Ph[r * N + i];
        Q[pp * N + i] = acc / (double)N;
      }
      fft3d(g, Q.data() + pp * N, +1);
    }
    double muF = 0.0;
    for (int v = 0; v < V; ++v) {
      if (p->noise != 0.0) or_noise_field(p->seed, (uint32_t)(step0 + s), (uint32_t)v, N, xi.data());
      for (int64_t i = 0; i < N; ++i) {
        double s2 = 0.0;
        for (int u = 0; u < V; ++u) s2 += e[u * N + i] * e[u * N + i];
        const double ev = e[v * N + i];
        double f = p->A * ev - p->B * ev * ev + p->C * ev * s2 + 2.0 * w * ev * R[v * N + i].real();
        if (p->noise != 0.0) f += p->noise * (double)xi[i];
        F[v * N + i] = f;
        muF += mu[v * N + i] * f;
      }
    }
    gL += -p->dt * muF;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
      double s2 = 0.0, t = 0.0;
      for (int u = 0; u < V; ++u) { s2 += e[u * N + i] * e[u * N + i]; t += e[u * N + i] * mu[u * N + i]; }
      for (int v = 0; v < V; ++v) {
        const double ev = e[v * N + i], m = mu[v * N + i];
        const double jt = (p->A - 2.0 * p->B * ev + p->C * s2 + 2.0 * w * R[v * N + i].real()) * m +
                          2.0 * p->C * ev * t + 4.0 * w * ev * Q[v * N + i].real();
        lam[v * N + i] = m - dtL * jt;
      }
    }
  }
  std::memcpy(g_eta0, lam.data(), VN * sizeof(double));
  for (int v = 0; v < V; ++v) g_par[v] = gs[v];
  g_par[V] = gL;
  g_par[V + 1] = gk;
}

static void mart_energy(const Grid& g, int V, const or_mart_params* p, const double* B, const double* eta,
                        double out[4]) {
  const int64_t N = g.N;
  double ech = 0, egr = 0, eel = 0;
  std::vector<cd> R;
  elastic_R(g, V, B, eta, R);
#pragma omp parallel for reduction(+ : ech, eel) schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    double s2 = 0, s3 = 0, el = 0;
    for (int v = 0; v < V; ++v) {
      double e = eta[v * N + i];
      s2 += e * e; s3 += e * e * e;
      el += e * e * R[v * N + i].real();
    }
    ech += 0.5 * p->A * s2 - p->B / 3.0 * s3 + 0.25 * p->C * s2 * s2;
    eel += 0.5 * p->w_el * el;
  }
  const KSq kk(g, p->kconv);
  std::vector<cd> eh(N);
  for (int v = 0; v < V; ++v) {
    for (int64_t i = 0; i < N; ++i) eh[i] = cd(eta[v * N + i], 0);
    fft3d(g, eh.data(), -1);
    double acc = 0;
#pragma omp parallel for reduction(+ : acc) schedule(static)
    for (int64_t i = 0; i < N; ++i) acc += kk(i) * std::norm(eh[i]);
    egr += 0.5 * p->kappa * acc / (double)N;
  }
  out[0] = ech; out[1] = egr; out[2] = eel; out[3] = ech + egr + eel;
}
extern "C" void or_mart_energy(int n, int V, const or_mart_params* p, const double* B, const double* eta,
                               double out[4]) {
  mart_energy(cube_grid(n, p->dx), V, p, B, eta, out);
}
extern "C" void or_mart_energy_box(const or_grid* g, int V, const or_mart_params* p, const double* B,
                                   const double* eta, double out[4]) {
  mart_energy(box_grid(g), V, p, B, eta, out);
}

// ---------------------------------------------------------------------------
// Precipitate: CH for c (k^4 implicit) + AC for eta (k^2 implicit)
//   mu_c (src/dynamics_precipitate.cu:118-127), F_v (:136-162),
//   c update (:197-206), eta update (:234-248).
// ---------------------------------------------------------------------------
extern "C" void or_prec_local(int V, const or_prec_params* p, double c, const double* eta, double* mu_c, double* F,
                              double* f) {
  const double* A = p->A;
  double s2 = 0;
  for (int v = 0; v < V; ++v) s2 += eta[v] * eta[v];
  if (mu_c) *mu_c = 0.5 * A[2] * s2 + A[1] * (c - p->c1);
  if (F) {
    for (int v = 0; v < V; ++v) {
      double e = eta[v], t2 = 0, t4 = 0, t22 = 0;
      for (int i = 0; i < V; ++i)
        if (i != v) { double q = eta[i] * eta[i]; t2 += q; t4 += q * q; }
      for (int i = 0; i < V; ++i)
        for (int j = 0; j < V; ++j)
          if (v != i && v != j && i != j) t22 += eta[i] * eta[i] * eta[j] * eta[j];
      F[v] = (2.0 * A[5] * e + 4.0 * A[6] * e * e * e) * t2 + 2.0 * A[6] * e * t4 + 2.0 * A[7] * e * t22 +
             A[2] * e * (c - p->c2) - A[3] * e * e * e + A[4] * e * e * e * e * e;
    }
  }
  if (f) {
    double v0 = 0.5 * A[1] * (c - p->c1) * (c - p->c1) + 0.5 * A[2] * (c - p->c2) * s2;
    for (int v = 0; v < V; ++v) {
      double q = eta[v] * eta[v];
      v0 += -0.25 * A[3] * q * q + A[4] / 6.0 * q * q * q;
    }
    for (int i = 0; i < V; ++i)
      for (int j = i + 1; j < V; ++j) {
        double a = eta[i] * eta[i], b = eta[j] * eta[j];
        v0 += A[5] * a * b + A[6] * (a * a * b + a * b * b);
        for (int k = j + 1; k < V; ++k) v0 += 2.0 * A[7] * a * b * eta[k] * eta[k];
      }
    *f = v0;
  }
}

static void prec_run(const Grid& g, int V, const or_prec_params* p, const double* Bc, double* c, double* eta,
                     int step0, int nsteps) {
  const int64_t N = g.N;
  const KSq kk(g, p->kconv);
  std::vector<double> mu(N), F((size_t)V * N);
  std::vector<float> xi(N);
  std::vector<cd> ch(N), mh(N), eh(N), fh(N);
  for (int s = 0; s < nsteps; ++s) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
      double e[16], Fl[16];
      for (int v = 0; v < V; ++v) e[v] = eta[v * N + i];
      or_prec_local(V, p, c[i], e, &mu[i], Fl, nullptr);
      for (int v = 0; v < V; ++v) F[v * N + i] = Fl[v];
    }
    // concentration
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) { ch[i] = cd(c[i], 0); mh[i] = cd(mu[i], 0); }
    fft3d(g, ch.data(), -1);
    fft3d(g, mh.data(), -1);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
      const double k2 = kk(i);
      cd mt = mh[i] + (p->xi * Bc[i] - p->S) * ch[i];
      ch[i] = (ch[i] - p->dt * p->M * k2 * mt) / (1.0 + p->dt * p->M * (p->S * k2 + p->kappa_c * k2 * k2)) /
              (double)N;
    }
    fft3d(g, ch.data(), +1);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) c[i] = ch[i].real();
    // order parameters
    for (int v = 0; v < V; ++v) {
      if (p->noise != 0.0) {
        or_noise_field(p->seed, (uint32_t)(step0 + s), (uint32_t)v, N, xi.data());
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; ++i) F[v * N + i] += p->noise * (double)xi[i];
      }
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < N; ++i) { eh[i] = cd(eta[v * N + i], 0); fh[i] = cd(F[v * N + i], 0); }
      fft3d(g, eh.data(), -1);
      fft3d(g, fh.data(), -1);
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < N; ++i)
        eh[i] = (eh[i] - p->dt * p->L * fh[i]) / (1.0 + p->dt * p->L * p->kappa_eta * kk(i)) / (double)N;
      fft3d(g, eh.data(), +1);
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < N; ++i) eta[v * N + i] = eh[i].real();
    }
  }
}
extern "C" void or_prec_run(int n, int V, const or_prec_params* p, const double* Bc, double* c, double* eta,
                            int step0, int nsteps) {
  prec_run(cube_grid(n, p->dx), V, p, Bc, c, eta, step0, nsteps);
}
extern "C" void or_prec_run_box(const or_grid* g, int V, const or_prec_params* p, const double* Bc, double* c,
                                double* eta, int step0, int nsteps) {
  prec_run(box_grid(g), V, p, Bc, c, eta, step0, nsteps);
}

// ---------------------------------------------------------------------------
// Reference rules: the plugins' own updates, transcribed.
//   Martensite (src/dynamics_mart.cu:181-417): explicit Euler
//     eta += dt (w_grad Grad + Chem + El + Dis + w_noise xi), then eta *= (1 - defect)
//     Grad  = sum_d (eta(x+e_d) + eta(x-e_d) - 2 eta(x)) / (2 dx) / 3           (:181-188)
//     Chem  = HEAD: s = sqrt|a1/(a2-a3)|, eta >= 0: -w_chem (s eta)(a1 - a2 (s eta)^2 + a3 sum_i (s eta_i)^2),
//                   eta < 0: 1000 w_chem a2 eta^2                              (:222-250 HEAD side)
//             master: -w_chem eta (a1 - a2 eta + a3 sum_i eta_i)                (:222-250 master side,
//                   whose eta <= 0 test compares a pointer and is never taken)
//     El    = -w_el R_p eta_p, R_p = IFFT[sum_q B_pq FFT(eta_q^2)]/N, all q       (:270-289 HEAD)
//     Dis   = -w_dis K_p eta_p, K_p = 25 sum_ij sigma_ij eps^p_ij (given)        (:305-331)
//   Precipitate (src/dynamics_precipitate.cu:118-289): the spectral rule of or_prec_run with
//     the eta^2 elastic source El_v = 2 R_v eta_v added as xi El_v to F_v (:170-195, :249),
//     no elasticity and no stabilisation in the c update (:197-206), and noise of amplitude
//     p->noise on both fields (the reference's 1e-4 weightnoise, :203, :244): eta_v uses
//     Philox field v, c uses field 0x8000.
// ---------------------------------------------------------------------------
static void mart_ref_run(const Grid& g, int V, const or_mart_ref_params* p, const double* B, double* eta,
                         const double* defect, const double* K, int step0, int nsteps) {
  const int64_t N = g.N, VN = (int64_t)V * N;
  std::vector<double> out((size_t)VN);
  std::vector<float> xi((size_t)VN, 0.0f);
  std::vector<cd> R;
  // 1 / (2 Delta_d) / 3 per axis (src/dynamics_mart.cu:181-188)
  const double gx = 1.0 / (2.0 * g.dx) / 3.0, gy = 1.0 / (2.0 * g.dy) / 3.0, gz = 1.0 / (2.0 * g.dz) / 3.0;
  const double sc = std::sqrt(std::fabs(p->a1 / (p->a2 - p->a3)));
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  for (int s = 0; s < nsteps; ++s) {
    elastic_R(g, V, B, eta, R);
    if (p->w_noise != 0.0)
      for (int v = 0; v < V; ++v) or_noise_field(p->seed, (uint32_t)(step0 + s), (uint32_t)v, N, xi.data() + v * N);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
      int x, y, z;
      g.coords(i, x, y, z);
      const int64_t ox = (int64_t)ny * nz, oy = nz;
      const int64_t xp = i + ((x + 1) % nx - x) * ox, xm = i + ((x + nx - 1) % nx - x) * ox;
      const int64_t yp = i + ((y + 1) % ny - y) * oy, ym = i + ((y + ny - 1) % ny - y) * oy;
      const int64_t zp = i + ((z + 1) % nz - z), zm = i + ((z + nz - 1) % nz - z);
      double S2 = 0.0, S1 = 0.0;
      for (int u = 0; u < V; ++u) {
        const double e = eta[u * N + i];
        S2 += (sc * e) * (sc * e);
        S1 += e;
      }
      for (int v = 0; v < V; ++v) {
        const double* ev = eta + v * N;
        const double e = ev[i];
        const double grad = (ev[xp] + ev[xm] - 2.0 * e) * gx + (ev[yp] + ev[ym] - 2.0 * e) * gy +
                            (ev[zp] + ev[zm] - 2.0 * e) * gz;
        double chem;
        if (p->chem_form == 0)
          chem = e >= 0.0 ? -p->w_chem * (sc * e) * (p->a1 - p->a2 * (sc * e) * (sc * e) + p->a3 * S2)
                          : 1000.0 * p->w_chem * p->a2 * e * e;
        else
          chem = -p->w_chem * e * (p->a1 - p->a2 * e + p->a3 * S1);
        const double el = -p->w_el * R[v * N + i].real() * e;
        const double dis = K ? -p->w_dis * K[v * N + i] * e : 0.0;
        double nv = e + p->dt * (p->w_grad * grad + chem + el + dis + p->w_noise * (double)xi[v * N + i]);
        if (defect) nv *= 1.0 - defect[i];
        out[v * N + i] = nv;
      }
    }
    std::memcpy(eta, out.data(), VN * sizeof(double));
  }
}
extern "C" void or_mart_ref_run(int n, int V, const or_mart_ref_params* p, const double* B, double* eta,
                                const double* defect, const double* K, int step0, int nsteps) {
  mart_ref_run(cube_grid(n, p->dx), V, p, B, eta, defect, K, step0, nsteps);
}
extern "C" void or_mart_ref_run_box(const or_grid* g, int V, const or_mart_ref_params* p, const double* B,
                                    double* eta, const double* defect, const double* K, int step0, int nsteps) {
  mart_ref_run(box_grid(g), V, p, B, eta, defect, K, step0, nsteps);
}

static void prec_ref_run(const Grid& g, int V, const or_prec_params* p, const double* B, double* c, double* eta,
                         int step0, int nsteps) {
  const int64_t N = g.N;
  const KSq kk(g, p->kconv);
  std::vector<double> mu(N), F((size_t)V * N);
  std::vector<float> xi(N);
  std::vector<cd> ch(N), mh(N), eh(N), fh(N), R;
  for (int s = 0; s < nsteps; ++s) {
    elastic_R(g, V, B, eta, R);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
      double e[16], Fl[16];
      for (int v = 0; v < V; ++v) e[v] = eta[v * N + i];
      or_prec_local(V, p, c[i], e, &mu[i], Fl, nullptr);
      for (int v = 0; v < V; ++v) F[v * N + i] = Fl[v] + p->xi * 2.0 * R[v * N + i].real() * e[v];
    }
    if (p->noise != 0.0) {
      or_noise_field(p->seed, (uint32_t)(step0 + s), 0x8000u, N, xi.data());
      for (int64_t i = 0; i < N; ++i) mu[i] += p->noise * (double)xi[i];
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) { ch[i] = cd(c[i], 0); mh[i] = cd(mu[i], 0); }
    fft3d(g, ch.data(), -1);
    fft3d(g, mh.data(), -1);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
      const double k2 = kk(i);
      ch[i] = (ch[i] - p->dt * p->M * k2 * mh[i]) / (1.0 + p->dt * p->M * p->kappa_c * k2 * k2) / (double)N;
    }
    fft3d(g, ch.data(), +1);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) c[i] = ch[i].real();
    for (int v = 0; v < V; ++v) {
      if (p->noise != 0.0) {
        or_noise_field(p->seed, (uint32_t)(step0 + s), (uint32_t)v, N, xi.data());
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; ++i) F[v * N + i] += p->noise * (double)xi[i];
      }
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < N; ++i) { eh[i] = cd(eta[v * N + i], 0); fh[i] = cd(F[v * N + i], 0); }
      fft3d(g, eh.data(), -1);
      fft3d(g, fh.data(), -1);
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < N; ++i)
        eh[i] = (eh[i] - p->dt * p->L * fh[i]) / (1.0 + p->dt * p->L * p->kappa_eta * kk(i)) / (double)N;
      fft3d(g, eh.data(), +1);
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < N; ++i) eta[v * N + i] = eh[i].real();
    }
  }
}
extern "C" void or_prec_ref_run(int n, int V, const or_prec_params* p, const double* B, double* c, double* eta,
                                int step0, int nsteps) {
  prec_ref_run(cube_grid(n, p->dx), V, p, B, c, eta, step0, nsteps);
}
extern "C" void or_prec_ref_run_box(const or_grid* g, int V, const or_prec_params* p, const double* B, double* c,
                                    double* eta, int step0, int nsteps) {
  prec_ref_run(box_grid(g), V, p, B, c, eta, step0, nsteps);
}

// Stress plugin (src/dynamics_stress.cu:126-185): stress_b = -xi IFFT[sum_v B_bv FFT(s_v)]/N for
// the 6 base strains b, s_v = eta_v^2 (source 2) or eta_v (source 1); defect = any |eta_v| > 0.9.
extern "C" void or_stress(int n, int V, const double* B, double xi, int source, const double* eta, double* stress,
                          double* defect) {
  const Grid g = cube_grid(n, 1.0);
  const int64_t N = g.N;
  std::vector<cd> S((size_t)V * N), R(N);
  for (int v = 0; v < V; ++v) {
    for (int64_t i = 0; i < N; ++i) {
      const double e = eta[v * N + i];
      S[v * N + i] = cd(source == 2 ? e * e : e, 0.0);
    }
    fft3d(g, S.data() + v * N, -1);
  }
  for (int b = 0; b < 6; ++b) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
      cd acc = 0;
      for (int v = 0; v < V; ++v) acc += B[(int64_t)(b * V + v) * N + i] * S[v * N + i];
      R[i] = acc / (double)N;
    }
    fft3d(g, R.data(), +1);
    for (int64_t i = 0; i < N; ++i) stress[b * N + i] = -xi * R[i].real();
  }
  if (defect)
    for (int64_t i = 0; i < N; ++i) {
      double d = 0.0;
      for (int v = 0; v < V; ++v)
        if (std::fabs(eta[v * N + i]) > 0.9) d = 1.0;
      defect[i] = d;
    }
}

static void prec_energy(const Grid& g, int V, const or_prec_params* p, const double* Bc, const double* c,
                        const double* eta, double out[4]) {
  const int64_t N = g.N;
  double ech = 0;
#pragma omp parallel for reduction(+ : ech) schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    double e[16], f;
    for (int v = 0; v < V; ++v) e[v] = eta[v * N + i];
    or_prec_local(V, p, c[i], e, nullptr, nullptr, &f);
    ech += f;
  }
  const KSq kk(g, p->kconv);
  std::vector<cd> h(N);
  double egr = 0, eel = 0;
  for (int v = -1; v < V; ++v) {
    const double* src = (v < 0) ? c : eta + (int64_t)v * N;
    for (int64_t i = 0; i < N; ++i) h[i] = cd(src[i], 0);
    fft3d(g, h.data(), -1);
    double acc = 0, accel = 0;
#pragma omp parallel for reduction(+ : acc, accel) schedule(static)
    for (int64_t i = 0; i < N; ++i) {
      const double a2 = std::norm(h[i]);
      acc += kk(i) * a2;
      if (v < 0) accel += Bc[i] * a2;
    }
    egr += 0.5 * (v < 0 ? p->kappa_c : p->kappa_eta) * acc / (double)N;
    if (v < 0) eel = 0.5 * p->xi * accel / (double)N;
  }
  out[0] = ech; out[1] = egr; out[2] = eel; out[3] = ech + egr + eel;
}
extern "C" void or_prec_energy(int n, int V, const or_prec_params* p, const double* Bc, const double* c,
                               const double* eta, double out[4]) {
  prec_energy(cube_grid(n, p->dx), V, p, Bc, c, eta, out);
}
extern "C" void or_prec_energy_box(const or_grid* g, int V, const or_prec_params* p, const double* Bc,
                                   const double* c, const double* eta, double out[4]) {
  prec_energy(box_grid(g), V, p, Bc, c, eta, out);
}

extern "C" double or_sum(const double* a, int64_t n) {
  double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int64_t i = 0; i < n; ++i) s += a[i];
  return s;
}
