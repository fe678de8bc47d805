// B-tensor builder (FP64), FP64 plain-path kernels and FP64 reductions.
#pragma once
#include "kernels_strided.h"

namespace gsk {

// ====================================================================== Khachaturyan B tensor
// B_pq(n) = C_ijkl e^p_ij e^q_kl - n_i s^p_ij Omega_jk s^q_kl n_l, s = C:e,
// Omega = (C_ikjl n_k n_l)^-1, or 0 when |det| < 1e-8; n = k/|k|, or 0 when
// |k| <= 5e-7 (src/gtensorb.cu:14-34, 110-182). Evaluated in FP64.
struct BArgs {
  double C[81];
  double s1[8 * 9];
  double s2[8 * 9];
  double cee[64];
  int V1, V2;
  double kfac;
};

// B at the wave vector (g0, g1, g2)
__device__ __forceinline__ void btensor_k(double g0, double g1, double g2, const BArgs& a, double* out) {
  const double mo = sqrt(g0 * g0 + g1 * g1 + g2 * g2);
  double u[3] = {0.0, 0.0, 0.0};
  if (mo > 5e-7) { u[0] = g0 / mo; u[1] = g1 / mo; u[2] = g2 / mo; }
  double m[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int l = 0; l < 3; ++l) acc += a.C[((i * 3 + k) * 3 + l) * 3 + j] * u[k] * u[l];
      m[i * 3 + j] = acc;
    }
  const double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                     m[2] * (m[3] * m[7] - m[4] * m[6]);
  double om[9];
  if (fabs(det) < 1e-8) {
#pragma unroll
    for (int i = 0; i < 9; ++i) om[i] = 0.0;
  } else {
    const double r = 1.0 / det;
    om[0] = (m[4] * m[8] - m[5] * m[7]) * r;
    om[1] = (m[2] * m[7] - m[1] * m[8]) * r;
    om[2] = (m[1] * m[5] - m[2] * m[4]) * r;
    om[3] = (m[5] * m[6] - m[3] * m[8]) * r;
    om[4] = (m[0] * m[8] - m[2] * m[6]) * r;
    om[5] = (m[2] * m[3] - m[0] * m[5]) * r;
    om[6] = (m[3] * m[7] - m[4] * m[6]) * r;
    om[7] = (m[1] * m[6] - m[0] * m[7]) * r;
    om[8] = (m[0] * m[4] - m[1] * m[3]) * r;
  }
  for (int p = 0; p < a.V1; ++p) {
    double av[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) av[j] = u[0] * a.s1[p * 9 + j] + u[1] * a.s1[p * 9 + 3 + j] + u[2] * a.s1[p * 9 + 6 + j];
    double aw[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) aw[k] = av[0] * om[k] + av[1] * om[3 + k] + av[2] * om[6 + k];
    for (int q = 0; q < a.V2; ++q) {
      double t = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k)
        t += aw[k] * (a.s2[q * 9 + k * 3] * u[0] + a.s2[q * 9 + k * 3 + 1] * u[1] + a.s2[q * 9 + k * 3 + 2] * u[2]);
      out[p * a.V2 + q] = a.cee[p * a.V2 + q] - t;
    }
  }
}

// B at the integer wave vector (mx, my, mz) of a cube: k = kfac m
__device__ __forceinline__ void btensor_m(int mx, int my, int mz, const BArgs& a, double* out) {
  btensor_k(a.kfac * mx, a.kfac * my, a.kfac * mz, a, out);
}

// B at FFT index (kx, ky, kz) with the reference's mapping m -> m - n for m >= n/2
// (src/gtensorb.cu:9-12): the Nyquist index n/2 is taken as -n/2.
__device__ __forceinline__ void btensor_at(int kx, int ky, int kz, int n, const BArgs& a, double* out) {
  btensor_m((kx < n / 2) ? kx : kx - n, (ky < n / 2) ? ky : ky - n, (kz < n / 2) ? kz : kz - n, a, out);
}

// B for products with half (R2C) spectra. The reference multiplies the full complex
// spectrum by B at index k and keeps the real part of the inverse transform, which equals
// multiplying by (B(m) + B(m'))/2, m' = m with its Nyquist components (index n/2) negated:
// the index -k of a Nyquist component is itself, and B is even. A half spectrum is
// implicitly Hermitian, so the half-spectrum kernels use this average (it differs from
// B(m) only on Nyquist planes, and only when B is not symmetric under a reflection).
__device__ __forceinline__ void btensor_half_spec(int kx, int ky, int kz, int n, const BArgs& a, double* out) {
  const int h = n / 2;
  const int mx = (kx < h) ? kx : kx - n, my = (ky < h) ? ky : ky - n, mz = (kz < h) ? kz : kz - n;
  btensor_m(mx, my, mz, a, out);
  if (kx == h || ky == h || kz == h) {
    double o2[64];
    btensor_m(kx == h ? -mx : mx, ky == h ? -my : my, kz == h ? -mz : mz, a, o2);
    for (int i = 0; i < a.V1 * a.V2; ++i) out[i] = 0.5 * (out[i] + o2[i]);
  }
}

// Full layout B[p][q][x][y][z], FP64.
template <int = 0>
__global__ void k_btensor_full(double* __restrict__ B, int n, BArgs a) {
  const long N3 = (long)n * n * n;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N3; i += (long)gridDim.x * blockDim.x) {
    const int x = (int)(i / ((long)n * n)), y = (int)((i / n) % n), z = (int)(i % n);
    double out[64];
    btensor_at(x, y, z, n, a, out);
    for (int pq = 0; pq < a.V1 * a.V2; ++pq) B[pq * N3 + i] = out[pq];
  }
}

// Box grid nx x ny x nz: B[pq][x][y][z] in T (FP64 evaluation), per-axis wave numbers
// k_d = kf_d * (m_d, or m_d - n_d for m_d >= n_d / 2) (src/gtensorb.cu:14-24). npq = V1 * V2 entries,
// or 1 (only B_00: the scalar tensor of the precipitate's c source).
template <typename T>
__global__ void k_btensor_box(T* __restrict__ B, int nx, int ny, int nz, double kfx, double kfy, double kfz, BArgs a,
                              int npq) {
  const long N3 = (long)nx * ny * nz;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N3; i += (long)gridDim.x * blockDim.x) {
    const int x = (int)(i / ((long)ny * nz)), y = (int)((i / nz) % ny), z = (int)(i % nz);
    double out[64];
    btensor_k(kfx * ((x < nx / 2) ? x : x - nx), kfy * ((y < ny / 2) ? y : y - ny), kfz * ((z < nz / 2) ? z : z - nz),
              a, out);
    for (int pq = 0; pq < npq; ++pq) B[pq * N3 + i] = (T)out[pq];
  }
}

// Tile-ordered FP32 layout for the fused x pass, packed p <= q (V1 == V2).
template <int = 0>
__global__ void k_btensor_tiles(float* __restrict__ Bt, Geo g, BArgs a) {
  const int n = g.n, W = g.W, T = n * W;
  const int V = a.V1, npq = V * (V + 1) / 2;
  const long total = (long)g.ntile * T;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int e = (int)(i % T);
    const long t = i / T;
    const STile s = xtile((int)t, g);
    const int c = e % W, kx = e / W;
    const int ky = s.ky0 + s.dky * c, kz = s.kz0 + s.dkz * c;
    double out[64];
    btensor_half_spec(kx, ky, kz, n, a, out);
    for (int p = 0; p < V; ++p)
      for (int q = p; q < V; ++q) {
        const int pq = p * V - (p * (p - 1)) / 2 + (q - p);
        Bt[(t * npq + pq) * T + e] = (float)out[p * V + q];
      }
  }
}

// ====================================================================== plain path
// One kernel per reference operation on full complex (C2C) spectra, in T = double (the FP64 plain
// path of cubes and boxes) or T = float (the FP32 path of box grids). Grids are nx x ny x nz with
// fields stored [f][x][y][z]; 3D transforms are three axis passes. Wave numbers per axis:
// k_d = kf_d * (m_d, or m_d - n_d for m_d >= n_d / 2).
template <typename T> struct Cx;
template <> struct Cx<float> { using T2 = float2; };
template <> struct Cx<double> { using T2 = double2; };
template <typename T> __device__ __forceinline__ typename Cx<T>::T2 cx(T a, T b) {
  typename Cx<T>::T2 r;
  r.x = a;
  r.y = b;
  return r;
}

// Quads: four consecutive voxels of one field, starting at a multiple of 4 (every box edge is a
// multiple of 16, so each field holds whole quads and a quad of T or of complex T is aligned to
// its own size). Each thread moves its quad with 16-byte accesses.
__device__ __forceinline__ void ld_quad(const float* p, float (&o)[4]) {
  const float4 q = *reinterpret_cast<const float4*>(p);
  o[0] = q.x, o[1] = q.y, o[2] = q.z, o[3] = q.w;
}
__device__ __forceinline__ void ld_quad(const double* p, double (&o)[4]) {
  const double2 a = reinterpret_cast<const double2*>(p)[0], b = reinterpret_cast<const double2*>(p)[1];
  o[0] = a.x, o[1] = a.y, o[2] = b.x, o[3] = b.y;
}
// a quad of complex values as real and imaginary parts
__device__ __forceinline__ void ld_quad(const float2* p, float (&re)[4], float (&im)[4]) {
  const float4 a = reinterpret_cast<const float4*>(p)[0], b = reinterpret_cast<const float4*>(p)[1];
  re[0] = a.x, im[0] = a.y, re[1] = a.z, im[1] = a.w, re[2] = b.x, im[2] = b.y, re[3] = b.z, im[3] = b.w;
}
__device__ __forceinline__ void ld_quad(const double2* p, double (&re)[4], double (&im)[4]) {
  for (int i = 0; i < 4; ++i) {
    const double2 a = p[i];
    re[i] = a.x, im[i] = a.y;
  }
}
__device__ __forceinline__ void st_quad(float2* p, const float (&re)[4], const float (&im)[4]) {
  reinterpret_cast<float4*>(p)[0] = make_float4(re[0], im[0], re[1], im[1]);
  reinterpret_cast<float4*>(p)[1] = make_float4(re[2], im[2], re[3], im[3]);
}
__device__ __forceinline__ void st_quad(double2* p, const double (&re)[4], const double (&im)[4]) {
  for (int i = 0; i < 4; ++i) p[i] = make_double2(re[i], im[i]);
}
__device__ __forceinline__ void st_quad(float* p, const float (&v)[4]) {
  *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
}
__device__ __forceinline__ void st_quad(double* p, const double (&v)[4]) {
  reinterpret_cast<double2*>(p)[0] = make_double2(v[0], v[1]);
  reinterpret_cast<double2*>(p)[1] = make_double2(v[2], v[3]);
}
constexpr int kPlainMaxV = 8;  // variants on the plain path (martensite 1..8, precipitate 1..7)

struct BoxDim {
  int nx, ny, nz;
  __device__ __forceinline__ void coords(long v, int& x, int& y, int& z) const {
    x = (int)(v / ((long)ny * nz));
    y = (int)((v / nz) % ny);
    z = (int)(v % nz);
  }
};

template <typename T> struct PlainKT {
  T A, B, C, two_wel, noise, dtL, dtLk, inv_n3;
  T kfx, kfy, kfz;
  BoxDim d;
  int V;
  uint32_t step;
  unsigned long long seed;
};
using PlainK = PlainKT<double>;

// |k|^2 at voxel (x, y, z) of the spectrum, and at voxel index v
template <typename T, class K> __device__ __forceinline__ T ksq_xyz(const K& k, int x, int y, int z) {
  const T kx = k.kfx * (T)((x < k.d.nx / 2) ? x : x - k.d.nx), ky = k.kfy * (T)((y < k.d.ny / 2) ? y : y - k.d.ny),
          kz = k.kfz * (T)((z < k.d.nz / 2) ? z : z - k.d.nz);
  return kx * kx + ky * ky + kz * kz;
}
template <typename T, class K> __device__ __forceinline__ T ksq_at(const K& k, long v) {
  int x, y, z;
  k.d.coords(v, x, y, z);
  return ksq_xyz<T>(k, x, y, z);
}

// Pairs: two real fields share one C2C transform as G = X + iY, and their transforms come back
// at k from G(k) and G(-k): X(k) = (G(k) + conj G(-k)) / 2, Y(k) = (G(k) - conj G(-k)) / 2i; the
// inverse of X + iY for two Hermitian spectra X, Y is x + iy. The box steps use this to halve
// the transforms of real fields.
// index of the voxel at -k (each coordinate m -> (n - m) mod n)
__device__ __forceinline__ long neg_voxel(const BoxDim& d, int x, int y, int z) {
  return ((long)(x ? d.nx - x : 0) * d.ny + (y ? d.ny - y : 0)) * d.nz + (z ? d.nz - z : 0);
}
template <typename T>
__device__ __forceinline__ void unpack2(typename Cx<T>::T2 g, typename Cx<T>::T2 gm, typename Cx<T>::T2& X,
                                        typename Cx<T>::T2& Y) {
  X = cx<T>(T(0.5) * (g.x + gm.x), T(0.5) * (g.y - gm.y));
  Y = cx<T>(T(0.5) * (g.y + gm.y), T(0.5) * (gm.x - g.x));
}
// X + iY
template <typename T>
__device__ __forceinline__ typename Cx<T>::T2 pack2(typename Cx<T>::T2 X, typename Cx<T>::T2 Y) {
  return cx<T>(X.x - Y.y, X.y + Y.x);
}

// contiguous rows (z axis): L rows of N per tile, rows r0 .. r0 + L - 1 of the buffer
template <typename T2, int N, bool FWD>
__global__ void __launch_bounds__(1024) p_zrows(T2* __restrict__ buf, int L, const T2* __restrict__ twN) {
  extern __shared__ __align__(16) unsigned char psm_raw[];
  T2* tw = reinterpret_cast<T2*>(psm_raw);
  T2* tile = tw + N;
  const long r0 = (long)blockIdx.x * L;
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = twN[i];
  for (int e = threadIdx.x; e < L * N; e += blockDim.x) {
    const int l = e / N, m = e % N;
    tile[l * (N + 1) + m] = buf[(r0 + l) * N + m];
  }
  __syncthreads();
  Lines Ls{L, L, N + 1, 0, 1};
  fft_lines<T2, N, FWD>(tile, Ls, tw);
  for (int e = threadIdx.x; e < L * N; e += blockDim.x) {
    const int l = e / N, m = e % N;
    buf[(r0 + l) * N + m] = tile[l * (N + 1) + m];
  }
}

// strided lines: block -> (outer, chunk of W columns out of nch chunks); element (pos, c) at
// outer*os + pos*ps + chunk*W + c
template <typename T2, int N, bool FWD>
__global__ void __launch_bounds__(1024) p_strided(T2* __restrict__ buf, int W, int nch, long os, long ps,
                                                 const T2* __restrict__ twN) {
  extern __shared__ __align__(16) unsigned char psm_raw[];
  T2* tw = reinterpret_cast<T2*>(psm_raw);
  T2* tile = tw + N;
  const long o = blockIdx.x / nch;
  const int ch = blockIdx.x % nch;
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = twN[i];
  T2* base = buf + o * os + (long)ch * W;
  for (int e = threadIdx.x; e < N * W; e += blockDim.x) tile[e] = base[(e / W) * ps + e % W];
  __syncthreads();
  Lines Ls{W, W, 1, 0, W};
  fft_lines<T2, N, FWD>(tile, Ls, tw);
  for (int e = threadIdx.x; e < N * W; e += blockDim.x) base[(e / W) * ps + e % W] = tile[e];
}

template <typename T>
__global__ void p_square(const T* __restrict__ eta, typename Cx<T>::T2* __restrict__ S, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const T e = eta[i];
    S[i] = cx<T>(e * e, T(0));
  }
}

template <typename T>
__global__ void p_load(const T* __restrict__ eta, typename Cx<T>::T2* __restrict__ E, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    E[i] = cx<T>(eta[i], T(0));
}

template <typename T>
__global__ void p_extract(const typename Cx<T>::T2* __restrict__ E, T* __restrict__ eta, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) eta[i] = E[i].x;
}

// ---- the martensite step with paired transforms (box grids). P = ceil(V / 2) pairs; field 2j
// of a pair is the real part, field 2j + 1 (when it exists) the imaginary part.
// S_j = eta_2j^2 + i eta_2j+1^2
template <typename T>
__global__ void p_square_pair(const T* __restrict__ eta, typename Cx<T>::T2* __restrict__ S, long N3, int V) {
  for (long g4 = blockIdx.x * (long)blockDim.x + threadIdx.x; g4 < N3 / 4; g4 += (long)gridDim.x * blockDim.x) {
    const long v0 = g4 * 4;
#pragma unroll
    for (int j = 0; j < kPlainMaxV / 2; ++j)
      if (2 * j < V) {
        T a[4], b[4] = {0, 0, 0, 0};
        ld_quad(eta + 2 * j * N3 + v0, a);
        if (2 * j + 1 < V) ld_quad(eta + (2 * j + 1) * N3 + v0, b);
        for (int i = 0; i < 4; ++i) a[i] *= a[i], b[i] *= b[i];
        st_quad(S + j * N3 + v0, a, b);
      }
  }
}

// R_p = sum_q B_pq S_q / N with S_q unpacked from the transformed pairs; R packed as pairs. V is a
// template argument so that the V^2 loads of B carry no guards and are all in flight together.
template <typename T, int V>
__global__ void __launch_bounds__(256) p_contract_pair(const typename Cx<T>::T2* __restrict__ S, typename Cx<T>::T2* __restrict__ R,
                                const T* __restrict__ B, long N3, PlainKT<T> k) {
  using T2 = typename Cx<T>::T2;
  constexpr int P = (V + 1) / 2;
  for (long v = blockIdx.x * (long)blockDim.x + threadIdx.x; v < N3; v += (long)gridDim.x * blockDim.x) {
    int x, y, z;
    k.d.coords(v, x, y, z);
    const long nv = neg_voxel(k.d, x, y, z);
    T2 g[P], gm[P], sq[2 * P], r[2 * P];
    T b[V][V];
#pragma unroll
    for (int j = 0; j < P; ++j) g[j] = S[j * N3 + v], gm[j] = S[j * N3 + nv];
#pragma unroll
    for (int p = 0; p < V; ++p)
#pragma unroll
      for (int q = 0; q < V; ++q) b[p][q] = B[(long)(p * V + q) * N3 + v];
#pragma unroll
    for (int j = 0; j < P; ++j) unpack2<T>(g[j], gm[j], sq[2 * j], sq[2 * j + 1]);
#pragma unroll
    for (int p = 0; p < V; ++p) {
      T ax = 0, ay = 0;
#pragma unroll
      for (int q = 0; q < V; ++q) {
        ax += b[p][q] * sq[q].x;
        ay += b[p][q] * sq[q].y;
      }
      r[p] = cx<T>(ax * k.inv_n3, ay * k.inv_n3);
    }
    if constexpr ((V & 1) != 0) r[V] = cx<T>(T(0), T(0));
#pragma unroll
    for (int j = 0; j < P; ++j) R[j * N3 + v] = pack2<T>(r[2 * j], r[2 * j + 1]);
  }
}

// G_p = F_p + i eta_p: F_p = A eta_p - B eta_p^2 + C eta_p sum_q eta_q^2 + 2 w_el eta_p R_p
// (+ noise xi_p, Philox field p, keyed by the quad), R_p from the packed pairs. One thread per quad
// of voxels; the variants stay in registers.
template <typename T>
__global__ void __launch_bounds__(256) p_force_pair(const T* __restrict__ eta, const typename Cx<T>::T2* __restrict__ R,
                             typename Cx<T>::T2* __restrict__ G, long N3, PlainKT<T> k) {
  for (long g4 = blockIdx.x * (long)blockDim.x + threadIdx.x; g4 < N3 / 4; g4 += (long)gridDim.x * blockDim.x) {
    const long v0 = g4 * 4;
    T e[kPlainMaxV][4], s2[4] = {0, 0, 0, 0}, r[2][4];
#pragma unroll
    for (int p = 0; p < kPlainMaxV; ++p)
      if (p < k.V) {
        ld_quad(eta + p * N3 + v0, e[p]);
        for (int i = 0; i < 4; ++i) s2[i] += e[p][i] * e[p][i];
      }
#pragma unroll
    for (int p = 0; p < kPlainMaxV; ++p)
      if (p < k.V) {
        if (!(p & 1)) ld_quad(R + (p >> 1) * N3 + v0, r[0], r[1]);
        T f[4];
        for (int i = 0; i < 4; ++i) {
          const T ev = e[p][i];
          f[i] = k.A * ev - k.B * ev * ev + k.C * ev * s2[i] + k.two_wel * ev * r[p & 1][i];
        }
        if (k.noise != T(0)) {
          float xi[4];
          normal4(k.seed, (uint32_t)g4, (uint32_t)p, k.step, xi);
          for (int i = 0; i < 4; ++i) f[i] += k.noise * (T)xi[i];
        }
        st_quad(G + p * N3 + v0, f, e[p]);
      }
  }
}

// eta_hat_p <- (eta_hat_p - dt L F_hat_p) / (1 + dt L kappa k^2) / N with both unpacked from G_p;
// the new spectra packed as pairs into H
template <typename T>
__global__ void p_update_pair(const typename Cx<T>::T2* __restrict__ G, typename Cx<T>::T2* __restrict__ H,
                              PlainKT<T> k) {
  using T2 = typename Cx<T>::T2;
  const long N3 = (long)k.d.nx * k.d.ny * k.d.nz;
  for (long v = blockIdx.x * (long)blockDim.x + threadIdx.x; v < N3; v += (long)gridDim.x * blockDim.x) {
    int x, y, z;
    k.d.coords(v, x, y, z);
    const long nv = neg_voxel(k.d, x, y, z);
    const T den = T(1) / (T(1) + k.dtLk * ksq_xyz<T>(k, x, y, z)) * k.inv_n3;
    T2 en[kPlainMaxV];
#pragma unroll
    for (int p = 0; p < kPlainMaxV; ++p)
      if (p < k.V) {
        T2 f, e;
        unpack2<T>(G[p * N3 + v], G[p * N3 + nv], f, e);
        en[p] = cx<T>((e.x - k.dtL * f.x) * den, (e.y - k.dtL * f.y) * den);
      }
#pragma unroll
    for (int j = 0; j < kPlainMaxV / 2; ++j)
      if (2 * j < k.V) H[j * N3 + v] = pack2<T>(en[2 * j], 2 * j + 1 < k.V ? en[2 * j + 1] : cx<T>(T(0), T(0)));
  }
}

// field f of nf real fields from the real (f even) or imaginary (f odd) part of pair f / 2
template <typename T>
__global__ void p_extract_pair(const typename Cx<T>::T2* __restrict__ H, T* __restrict__ out, long N3, int nf) {
  for (long g4 = blockIdx.x * (long)blockDim.x + threadIdx.x; g4 < N3 / 4; g4 += (long)gridDim.x * blockDim.x) {
    const long v0 = g4 * 4;
#pragma unroll
    for (int j = 0; j < kPlainMaxV / 2; ++j)
      if (2 * j < nf) {
        T a[4], b[4];
        ld_quad(H + j * N3 + v0, a, b);
        st_quad(out + 2 * j * N3 + v0, a);
        if (2 * j + 1 < nf) st_quad(out + (2 * j + 1) * N3 + v0, b);
      }
  }
}

// ---- precipitate on the plain path (src/dynamics_precipitate.cu:118-275), one kernel per
// operation. Fields: [c, eta_1..eta_V], F = V + 1. The spectral rule has the c-sourced elasticity
// and the stabilisation S; the reference rule (Rel != null) adds xi 2 R_v eta_v to F_v
// (src/dynamics_precipitate.cu:170-195, :249) and noise on c (Philox field 0x8000).
template <typename T> struct PlainPKT {
  T A[8];
  T c1, c2, dtM, dtMkc, dtL, dtLke, xi, S, noise, inv_n3;
  T cbar;  // c = cbar + the stored field 0 (the FP32 box path stores the fluctuation; 0 otherwise)
  T kfx, kfy, kfz;
  BoxDim d;
  int V;
  int celastic;  // 1: the spectral rule's xi B(k) c_hat term in the concentration update
  uint32_t step;
  unsigned long long seed;
};
using PlainPK = PlainPKT<double>;

// mu_c = A2/2 sum eta^2 + A1 (c - c1) (:118-127) and F_v (:136-162) + noise a xi_v (Philox field v)
// of every voxel into Fh; the fields themselves into E (both as complex with zero imaginary part).
// Rel (reference rule, may be null): the elastic R_v of the eta^2 source as complex spectra whose
// real parts hold R_v. Each force shares a complex field with its own field (Fh_f = F_f + i
// field_f, one C2C transform for both), and Rel holds R as packed pairs.
template <typename T>
__global__ void __launch_bounds__(256) p_prec_local(const T* __restrict__ fld, typename Cx<T>::T2* __restrict__ Fh,
                                                    long N3, PlainPKT<T> k,
                                                    const typename Cx<T>::T2* __restrict__ Rel) {
  const int V = k.V;
  for (long g4 = blockIdx.x * (long)blockDim.x + threadIdx.x; g4 < N3 / 4; g4 += (long)gridDim.x * blockDim.x) {
    const long v0 = g4 * 4;
    T c0[4], e[kPlainMaxV][4], s2[4] = {0, 0, 0, 0}, mu[4];
    ld_quad(fld + v0, c0);
#pragma unroll
    for (int a = 0; a < kPlainMaxV; ++a)
      if (a < V) {
        ld_quad(fld + (1 + a) * N3 + v0, e[a]);
        for (int i = 0; i < 4; ++i) s2[i] += e[a][i] * e[a][i];
      }
    for (int i = 0; i < 4; ++i) mu[i] = T(0.5) * k.A[2] * s2[i] + k.A[1] * (c0[i] + k.cbar - k.c1);
    if (Rel && k.noise != T(0)) {
      float xi[4];
      normal4(k.seed, (uint32_t)g4, 0x8000u, k.step, xi);
      for (int i = 0; i < 4; ++i) mu[i] += k.noise * (T)xi[i];
    }
    st_quad(Fh + v0, mu, c0);
    T rp[2][4];
#pragma unroll
    for (int a = 0; a < kPlainMaxV; ++a)
      if (a < V) {
        T r[4], f[4];
        if (Rel) {
          if (!(a & 1)) ld_quad(Rel + (a >> 1) * N3 + v0, rp[0], rp[1]);
          for (int i = 0; i < 4; ++i) r[i] = rp[a & 1][i];
        }
        for (int i = 0; i < 4; ++i) {
          T t2 = 0, t4 = 0;
#pragma unroll
          for (int b = 0; b < kPlainMaxV; ++b)
            if (b < V && b != a) {
              const T q = e[b][i] * e[b][i];
              t2 += q;
              t4 += q * q;
            }
          const T c = c0[i] + k.cbar, et = e[a][i], e2 = et * et;
          // ordered pairs i != j, both != a: (sum q)^2 - sum q^2
          f[i] = (T(2) * k.A[5] * et + T(4) * k.A[6] * e2 * et) * t2 + T(2) * k.A[6] * et * t4 +
                 T(2) * k.A[7] * et * (t2 * t2 - t4) + k.A[2] * et * (c - k.c2) - k.A[3] * e2 * et +
                 k.A[4] * e2 * e2 * et;
          if (Rel) f[i] += k.xi * T(2) * r[i] * et;
        }
        if (k.noise != T(0)) {
          float xi[4];
          normal4(k.seed, (uint32_t)g4, (uint32_t)a, k.step, xi);
          for (int i = 0; i < 4; ++i) f[i] += k.noise * (T)xi[i];
        }
        st_quad(Fh + (1 + a) * N3 + v0, f, e[a]);
      }
  }
}

// Semi-implicit updates of the transformed fields (with the 1/N of the inverse transform):
//   c_hat   <- (c_hat - dt M k^2 (mu_hat + (xi B(k) - S) c_hat)) / (1 + dt M (S k^2 + kappa_c k^4))
//   eta_hat <- (eta_hat - dt L F_hat) / (1 + dt L kappa_eta k^2)
// (celastic = 0: no xi B(k) term; the reference rule also has S = 0)
// c_hat, mu_hat and eta_hat_v, F_hat_v unpacked from G_f = F_f + i field_f (the output of
// p_prec_local after its transform); the new spectra packed as pairs into H
template <typename T>
__global__ void p_prec_update_pair(const typename Cx<T>::T2* __restrict__ G, typename Cx<T>::T2* __restrict__ H,
                                   const T* __restrict__ Bc, PlainPKT<T> k) {
  using T2 = typename Cx<T>::T2;
  const long N3 = (long)k.d.nx * k.d.ny * k.d.nz;
  const int F = k.V + 1;
  for (long v = blockIdx.x * (long)blockDim.x + threadIdx.x; v < N3; v += (long)gridDim.x * blockDim.x) {
    int x, y, z;
    k.d.coords(v, x, y, z);
    const long nv = neg_voxel(k.d, x, y, z);
    const T k2 = ksq_xyz<T>(k, x, y, z);
    T2 en[kPlainMaxV];
#pragma unroll
    for (int f = 0; f < kPlainMaxV; ++f)
      if (f < F) {
        T2 fh, e;
        unpack2<T>(G[f * N3 + v], G[f * N3 + nv], fh, e);
        if (f == 0) {
          const T a = (k.celastic ? k.xi * Bc[v] : T(0)) - k.S;
          const T mx = fh.x + a * e.x, my = fh.y + a * e.y;
          const T den = k.inv_n3 / (T(1) + k.dtM * k.S * k2 + k.dtMkc * k2 * k2);
          en[f] = cx<T>((e.x - k.dtM * k2 * mx) * den, (e.y - k.dtM * k2 * my) * den);
        } else {
          const T den = k.inv_n3 / (T(1) + k.dtLke * k2);
          en[f] = cx<T>((e.x - k.dtL * fh.x) * den, (e.y - k.dtL * fh.y) * den);
        }
      }
#pragma unroll
    for (int j = 0; j < kPlainMaxV / 2; ++j)
      if (2 * j < F) H[j * N3 + v] = pack2<T>(en[2 * j], 2 * j + 1 < F ? en[2 * j + 1] : cx<T>(T(0), T(0)));
  }
}

// ====================================================================== FP64 reductions
// Two-stage deterministic sums: each block writes a partial, a single block adds them.
__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < (int)((blockDim.x + 63) / 64); ++i) r += sh[i];
  }
  __syncthreads();
  return r;
}

template <int = 0>
__global__ void r_finish(const double* __restrict__ part, int np, int nout, double* __restrict__ out) {
  // part[b*nout + j]
  __shared__ double sh[16];
  for (int j = 0; j < nout; ++j) {
    double v = 0.0;
    for (int b = threadIdx.x; b < np; b += blockDim.x) v += part[(long)b * nout + j];
    v = block_sum(v, sh);
    if (threadIdx.x == 0) out[j] = v;
  }
}

// Real-space terms: martensite chemical energy, sum eta^2, per-variant sums.
// out slots: 0 chem, 1 sum eta^2, 2.. 2+V per-variant sums
template <typename T>
__global__ void r_mart_real(const T* __restrict__ eta, long N3, int V, double A, double B, double C,
                            double* __restrict__ part) {
  __shared__ double sh[16];
  double chem = 0.0, s2t = 0.0, sv[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N3; i += (long)gridDim.x * blockDim.x) {
    double s2 = 0.0, s3 = 0.0;
    for (int v = 0; v < V; ++v) {
      const double e = (double)eta[v * N3 + i];
      s2 += e * e; s3 += e * e * e; sv[v] += e;
    }
    chem += 0.5 * A * s2 - B / 3.0 * s3 + 0.25 * C * s2 * s2;
    s2t += s2;
  }
  const int nout = 2 + V;
  double r = block_sum(chem, sh);
  if (threadIdx.x == 0) part[blockIdx.x * nout + 0] = r;
  r = block_sum(s2t, sh);
  if (threadIdx.x == 0) part[blockIdx.x * nout + 1] = r;
  for (int v = 0; v < V; ++v) {
    r = block_sum(sv[v], sh);
    if (threadIdx.x == 0) part[blockIdx.x * nout + 2 + v] = r;
  }
}

// Precipitate real-space: chemical energy (exact antiderivative of the local forces),
// sum of dc, per-variant sums. c = cbar + dc.
__device__ __forceinline__ double prec_f(double c, const double* e, int V, const double* A, double c1, double c2) {
  double s2 = 0.0;
  for (int v = 0; v < V; ++v) s2 += e[v] * e[v];
  double f = 0.5 * A[1] * (c - c1) * (c - c1) + 0.5 * A[2] * (c - c2) * s2;
  for (int v = 0; v < V; ++v) {
    const double q = e[v] * e[v];
    f += -0.25 * A[3] * q * q + A[4] / 6.0 * q * q * q;
  }
  for (int i = 0; i < V; ++i)
    for (int j = i + 1; j < V; ++j) {
      const double a = e[i] * e[i], b = e[j] * e[j];
      f += A[5] * a * b + A[6] * (a * a * b + a * b * b);
      for (int k = j + 1; k < V; ++k) f += 2.0 * A[7] * a * b * e[k] * e[k];
    }
  return f;
}

struct PrecE { double A[8]; double c1, c2, cbar; int V; };

template <typename T = float>
__global__ void r_prec_real(const T* __restrict__ fld, long N3, PrecE pe, double* __restrict__ part) {
  __shared__ double sh[16];
  double chem = 0.0, sdc = 0.0, sv[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int V = pe.V;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N3; i += (long)gridDim.x * blockDim.x) {
    double e[8];
    const double dc = (double)fld[i];
    for (int v = 0; v < V; ++v) { e[v] = (double)fld[(v + 1) * N3 + i]; sv[v] += e[v]; }
    chem += prec_f(pe.cbar + dc, e, V, pe.A, pe.c1, pe.c2);
    sdc += dc;
  }
  const int nout = 2 + V;
  double r = block_sum(chem, sh);
  if (threadIdx.x == 0) part[blockIdx.x * nout + 0] = r;
  r = block_sum(sdc, sh);
  if (threadIdx.x == 0) part[blockIdx.x * nout + 1] = r;
  for (int v = 0; v < V; ++v) {
    r = block_sum(sv[v], sh);
    if (threadIdx.x == 0) part[blockIdx.x * nout + 2 + v] = r;
  }
}

// Gradient energy from the tile-ordered fused state: sum_k w(kz) k^2 |X(k)|^2 per field
// (w = 1 for kz in {0, n/2}, 2 otherwise); optional k^2 exponent for Cahn-Hilliard is not needed.
template <int = 0>
__global__ void r_state_grad(const float2* __restrict__ state, Geo g, int F, float kfac, double* __restrict__ part) {
  __shared__ double sh[16];
  const int n = g.n, W = g.W, T = n * W;
  const long total = (long)g.ntile * F * T;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int e = (int)(i % T);
    const long tf = i / T;
    const int f = (int)(tf % F);
    const long t = tf / F;
    const STile s = xtile((int)t, g);
    const int c = e % W, kx = e / W;
    const int ky = s.ky0 + s.dky * c, kz = s.kz0 + s.dkz * c;
    const double kxv = (double)kfac * ((kx < n / 2) ? kx : kx - n), kyv = (double)kfac * ((ky < n / 2) ? ky : ky - n),
                 kzv = (double)kfac * kz;
    const double w = (kz == 0 || kz == g.M) ? 1.0 : 2.0;
    const float2 v = state[i];
    acc[f] += w * (kxv * kxv + kyv * kyv + kzv * kzv) * ((double)v.x * v.x + (double)v.y * v.y);
  }
  for (int f = 0; f < F; ++f) {
    double r = block_sum(acc[f], sh);
    if (threadIdx.x == 0) part[blockIdx.x * F + f] = r;
  }
}

// Precipitate elastic energy sum_k w B(k) |dc_hat(k)|^2 from the state (field 0).
template <int = 0>
__global__ void r_state_el_prec(const float2* __restrict__ state, const float* __restrict__ Bt, Geo g, int F,
                                double* __restrict__ part) {
  __shared__ double sh[16];
  const int W = g.W, T = g.n * W;
  const long total = (long)g.ntile * T;
  double acc = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int e = (int)(i % T);
    const long t = i / T;
    const STile st = xtile((int)t, g);
    const int kz = st.kz0 + st.dkz * (e % W);
    const double w = (kz == 0 || kz == g.M) ? 1.0 : 2.0;
    const float2 s = state[(t * F + 0) * T + e];
    acc += w * (double)Bt[t * T + e] * ((double)s.x * s.x + (double)s.y * s.y);
  }
  double r = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// Martensite elastic energy sum_k w sum_pq B_pq Re(conj(S_p) S_q) after the forward x transform of S.
template <int N>
__global__ void __launch_bounds__(xy_threads<N>()) r_xenergy_mart(const float2* __restrict__ kbuf, const float* __restrict__ Bt,
                                                      Geo g, int F, const float2* __restrict__ twN,
                                                      double* __restrict__ part) {
  extern __shared__ float2 smem[];
  __shared__ double sh[16];
  float2* tw = smem;
  float2* tiles = smem + N;
  const int W = g.W, T = N * W;
  const int t = blockIdx.x;
  const STile st = xtile(t, g);
  load_tw<N>(tw, twN);
  for (int f = 0; f < F; ++f) {
    const TileIO io = tio(const_cast<float2*>(kbuf) + f * g.kfs, st);
    for (int e = threadIdx.x; e < T; e += blockDim.x) tiles[f * T + e] = io.get(e % W, e / W);
  }
  __syncthreads();
  Lines Ls{W, W, 1, 0, W};
  for (int f = 0; f < F; ++f) fft_lines<float2, N, true>(tiles + f * T, Ls, tw);
  const int npq = F * (F + 1) / 2;
  const float* Bb = Bt + (long)t * npq * T;
  double acc = 0.0;
  for (int e = threadIdx.x; e < T; e += blockDim.x) {
    const int kz = st.kz0 + st.dkz * (e % W);
    const double w = (kz == 0 || kz == g.M) ? 1.0 : 2.0;
    double s = 0.0;
    for (int p = 0; p < F; ++p)
      for (int q = 0; q < F; ++q) {
        const int a = p < q ? p : q, b = p < q ? q : p;
        const int pq = a * F - (a * (a - 1)) / 2 + (b - a);
        const float2 sp = tiles[p * T + e], sq = tiles[q * T + e];
        s += (double)Bb[(long)pq * T + e] * ((double)sp.x * sq.x + (double)sp.y * sq.y);
      }
    acc += w * s;
  }
  double r = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// Same sum with the transformed tiles held in registers (V <= 4).
template <int N, int V>
__global__ void __launch_bounds__(xy_threads<N>()) r_xenergy_mart_r(float2* __restrict__ kbuf, const float* __restrict__ Bt,
                                                        Geo g, const float2* __restrict__ twN,
                                                        double* __restrict__ part) {
  constexpr int RL = edge_radix<N>();
  constexpr int NPQ = V * (V + 1) / 2;
  extern __shared__ float2 smem[];
  __shared__ double sh[16];
  float2* tw = smem;
  float2* tile = smem + N;
  const int W = g.W, T = N * W;
  const int t = blockIdx.x;
  const STile st = xtile(t, g);
  load_tw<N>(tw, twN);
  __syncthreads();
  Lines Ls{W, W, 1, 0, W};
  float2 S[V][RL];
#pragma unroll
  for (int f = 0; f < V; ++f) {
    const TileIO io = tio(kbuf + f * g.kfs, st);
    fft_g2r<float2, N>(tile, Ls, tw, [&](int c, int kx) { return io.get(c, kx); }, S[f]);
  }
  const int b = threadIdx.x;
  double acc = 0.0;
  const int kz = st.kz0 + st.dkz * (b % W);
  if (b < W * (N / RL)) {
    const double w = (kz == 0 || kz == g.M) ? 1.0 : 2.0;
    const float* Bb = Bt + (long)t * NPQ * T + b;
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      const int e = r * (N / RL) * W;
      double s = 0.0;
#pragma unroll
      for (int p = 0; p < V; ++p)
#pragma unroll
        for (int q = 0; q < V; ++q) {
          const int lo = p < q ? p : q, hi = p < q ? q : p;
          const int pq = lo * V - (lo * (lo - 1)) / 2 + (hi - lo);
          s += (double)Bb[(long)pq * T + e] *
               ((double)S[p][r].x * S[q][r].x + (double)S[p][r].y * S[q][r].y);
        }
      acc += w * s;
    }
  }
  double rr = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = rr;
}

// Plain-path energy pieces on full complex spectra: gradient sum k^2 |E|^2, elastic B_pq Re(S_p* S_q)
// (FP64 sums of T values).
template <typename T>
__global__ void r_plain_kspace(const typename Cx<T>::T2* __restrict__ E, const typename Cx<T>::T2* __restrict__ S,
                               const T* __restrict__ B, PlainKT<T> k, double* __restrict__ part) {
  __shared__ double sh[16];
  const long N3 = (long)k.d.nx * k.d.ny * k.d.nz;
  const int V = k.V;
  double gr = 0.0, el = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N3; i += (long)gridDim.x * blockDim.x) {
    const double k2 = (double)ksq_at<T>(k, i);
    for (int p = 0; p < V; ++p) {
      const typename Cx<T>::T2 e = E[p * N3 + i];
      gr += k2 * ((double)e.x * e.x + (double)e.y * e.y);
      for (int q = 0; q < V; ++q) {
        const typename Cx<T>::T2 sp = S[p * N3 + i], sq = S[q * N3 + i];
        el += (double)B[(long)(p * V + q) * N3 + i] * ((double)sp.x * sq.x + (double)sp.y * sq.y);
      }
    }
  }
  double r = block_sum(gr, sh);
  if (threadIdx.x == 0) part[blockIdx.x * 2] = r;
  r = block_sum(el, sh);
  if (threadIdx.x == 0) part[blockIdx.x * 2 + 1] = r;
}

// Precipitate, plain path: sum_k k^2 |c_hat|^2, sum_k k^2 sum_v |eta_hat_v|^2 and
// sum_k B(k) |c_hat|^2 over the full transformed fields E = [c, eta_1..eta_V].
template <typename T>
__global__ void r_plain_kspace_prec(const typename Cx<T>::T2* __restrict__ E, const T* __restrict__ Bc,
                                    PlainPKT<T> k, double* __restrict__ part) {
  __shared__ double sh[16];
  const long N3 = (long)k.d.nx * k.d.ny * k.d.nz;
  double gc = 0.0, ge = 0.0, el = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N3; i += (long)gridDim.x * blockDim.x) {
    const double k2 = (double)ksq_at<T>(k, i);
    const typename Cx<T>::T2 c = E[i];
    const double a2 = (double)c.x * c.x + (double)c.y * c.y;
    gc += k2 * a2;
    if (Bc) el += (double)Bc[i] * a2;
    for (int v = 1; v <= k.V; ++v) {
      const typename Cx<T>::T2 e = E[v * N3 + i];
      ge += k2 * ((double)e.x * e.x + (double)e.y * e.y);
    }
  }
  double r = block_sum(gc, sh);
  if (threadIdx.x == 0) part[blockIdx.x * 3] = r;
  r = block_sum(ge, sh);
  if (threadIdx.x == 0) part[blockIdx.x * 3 + 1] = r;
  r = block_sum(el, sh);
  if (threadIdx.x == 0) part[blockIdx.x * 3 + 2] = r;
}

// Device Philox normal field (for bit-exact host comparison).
template <int = 0>
__global__ void k_noise(float* __restrict__ out, long nvox, unsigned long long seed, uint32_t step, uint32_t field) {
  for (long g4 = blockIdx.x * (long)blockDim.x + threadIdx.x; g4 < nvox / 4; g4 += (long)gridDim.x * blockDim.x) {
    float xi[4];
    normal4(seed, (uint32_t)g4, field, step, xi);
    *reinterpret_cast<float4*>(out + 4 * g4) = make_float4(xi[0], xi[1], xi[2], xi[3]);
  }
}

// Philox uniform initial field lo + (hi - lo) * u, rounded to fp32 from fp64.
template <int = 0>
__global__ void k_uniform(float* __restrict__ out, long nvox, unsigned long long seed, uint32_t field, double lo,
                          double hi) {
  for (long g4 = blockIdx.x * (long)blockDim.x + threadIdx.x; g4 < nvox / 4; g4 += (long)gridDim.x * blockDim.x) {
    uint32_t u[4];
    uniform4_bits(seed, (uint32_t)g4, field, u);
    float4 r;
    r.x = (float)(lo + (hi - lo) * ((double)u[0] * (1.0 / 16777216.0)));
    r.y = (float)(lo + (hi - lo) * ((double)u[1] * (1.0 / 16777216.0)));
    r.z = (float)(lo + (hi - lo) * ((double)u[2] * (1.0 / 16777216.0)));
    r.w = (float)(lo + (hi - lo) * ((double)u[3] * (1.0 / 16777216.0)));
    *reinterpret_cast<float4*>(out + 4 * g4) = r;
  }
}

}  // namespace gsk
