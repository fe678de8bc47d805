// FP32 fused spectral step: shared geometry, physics constants and helpers.
//
// Field layout in HBM (matches the reference's Data<Real> [v][x][y][z], z
// fastest, include/datamain.th:70-76):
//   real fields     f * n^3 + (x*n + y)*n + z                       (float)
//   half spectrum   field f at f * kfs (float2), row r = kx*n + ky:
//                     kz <  n/2   r*(n/2) + kz        (rows of exactly n/2 values, so every
//                                                      row is a whole number of 128-byte lines)
//                     kz == n/2   nq + r, nq = n*n*n/2 (the Nyquist plane, contiguous)
//   x-pass state    tile-ordered: ((t*F + f)*n + kx)*W + c, t = x-pass tile (kernels_strided.h)
//   B tensor        tile-ordered: ((t*npq + pq)*n + kx)*W + c, packed p<=q
//
// Transform convention: forward R2C then C2C along y and x (unnormalised);
// the inverse C2C along x and y then C2R are unnormalised too, and the single
// 1/n^3 is applied in the x-pass between the two directions.
#pragma once
#include "fft_core.h"
#include "philox.h"

namespace gsk {

struct Geo {
  int n, M;          // edge, n/2 (also the row pitch of the half spectrum)
  int L;             // rows per z-pass tile
  int W, nchm;       // x/y-pass tile width and number of full kz chunks (n/2 = nchm*W)
  int ntile;         // x/y-pass tiles: n*nchm main tiles + n/W Nyquist tiles
  long long nvox;    // n^3
  long long kfs;     // complex elements per k-space field (>= n*n*(n/2+1), a multiple of 16)
  long long nq;      // offset of the Nyquist plane (kz = n/2) inside a field
  int frev, trev;    // visit fields / tiles of a per-field launch in reverse order
};

// Field and tile handled by this work group of a per-field launch (grid = tiles x fields).
// Consecutive launches alternate the visiting order so a kernel starts on the data its
// predecessor wrote last, which is the most likely to still sit in the last-level cache.
__device__ __forceinline__ int geo_field(const Geo& g) {
  return g.frev ? (int)(gridDim.y - 1 - blockIdx.y) : (int)blockIdx.y;
}
__device__ __forceinline__ int geo_tile(const Geo& g) {
  return g.trev ? (int)(gridDim.x - 1 - blockIdx.x) : (int)blockIdx.x;
}

// Index of half-spectrum element (field f, row r = kx*n + ky, kz) in the k-space buffer.
__host__ __device__ __forceinline__ long long kidx(const Geo& g, int f, long long r, int kz) {
  return f * g.kfs + (kz < g.M ? r * g.M + kz : g.nq + r);
}

// Physics constants for one launch (plain values, by value in kernel args).
struct MartK {
  float A, B, C;     // Landau
  float two_wel;     // 2 * elastic weight
  float noise;       // noise amplitude
  float dtL, dtLk;   // dt*L, dt*L*kappa
  float kfac;        // wave-number factor per index (2 pi/(n dx) or 2*3.14*dx/n)
  float inv_n3;
  int F;             // number of variants
  int npq;
  uint32_t step;
  const uint32_t* stepp;  // device step counter (read instead of `step` when set)
  unsigned long long seed;
};

struct PrecK {
  float A1, A2, A3, A4, A5, A6, A7;
  float c1, c2;
  float cbar;        // mean concentration (binary32 value kept in fp64 on the host)
  float k0;          // k = 0 mode of the concentration fluctuation
  float noise;
  float dtM, dtMkc;  // dt*M, dt*M*kappa_c
  float dtL, dtLke;  // dt*L, dt*L*kappa_eta
  float xi;          // elastic weight
  float S;           // Cahn-Hilliard stabilisation (implicit S k^2)
  int celastic;      // 1: c-sourced elasticity in the concentration update
  float kfac, inv_n3;
  int V;             // number of order parameters (fields = V + 1)
  uint32_t step;
  const uint32_t* stepp;  // device step counter (read instead of `step` when set)
  uint32_t* adv;          // device step counter advanced by the x update (fused step), or null
  unsigned long long seed;
};

// One member of a batched ensemble launch (grid.z = member): its buffers and constants. The
// batched kernels read entry blockIdx.z instead of their pointer and constant arguments.
struct MemberPtrs {
  float2* kbuf;
  float2* state;
  float* fld;
  const float* Bt;
  uint32_t* dstep;
  MartK mk;
};

// Step index of the noise counter: the device counter when the launch carries one
// (it is advanced on the device at the end of every fused step, so a captured graph
// replays with fresh noise), else the value passed at launch.
__device__ __forceinline__ uint32_t step_of(uint32_t step, const uint32_t* stepp) { return stepp ? *stepp : step; }

__device__ __forceinline__ float kcomp(int m, int n, float kfac) { return kfac * (float)((m < n / 2) ? m : m - n); }

// --------------------------------------------------------------- twiddles
// tw[m] = exp(-2 pi i m/N) for m < N  (N-point FFT)
// rt[k] = exp(-2 pi i k/n) for k < n/2 (R2C/C2R split)
template <int N>
__device__ __forceinline__ void load_tw(float2* dst, const float2* __restrict__ src) {
  for (int i = threadIdx.x; i < N; i += blockDim.x) dst[i] = src[i];
}

}  // namespace gsk
