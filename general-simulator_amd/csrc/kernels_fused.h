// FP32 fused spectral step kernels.
//
// Field layout in HBM (matches the reference's Data<Real> [v][x][y][z], z
// fastest, include/datamain.th:70-76):
//   real fields     f * n^3 + (x*n + y)*n + z                       (float)
//   half spectrum   f * n*n*Pz + (kx*n + ky)*Pz + kz, kz in [0, n/2] (float2)
//                   Pz = n/2 + 16 so every row starts on a 128-byte line.
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
  int n, M, Pz;      // edge, n/2, padded row pitch
  int L;             // rows per z-pass tile
  int W, nchm;       // x/y-pass tile width and number of full kz chunks (n/2 = nchm*W)
  int ntile;         // x/y-pass tiles: n*nchm main tiles + n/W Nyquist tiles
  long long nvox;    // n^3
  long long kfs;     // complex elements per k-space field
};

// Index of half-spectrum element (field f, row r = kx*n + ky, kz) in the k-space buffer.
__host__ __device__ __forceinline__ long long kidx(const Geo& g, int f, long long r, int kz) {
  return f * g.kfs + r * g.Pz + kz;
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
  float kfac, inv_n3;
  int V;             // number of order parameters (fields = V + 1)
  uint32_t step;
  unsigned long long seed;
};

__device__ __forceinline__ float kcomp(int m, int n, float kfac) { return kfac * (float)((m < n / 2) ? m : m - n); }

// --------------------------------------------------------------- twiddles
// tw[m] = exp(-2 pi i m/N) for m < N  (N-point FFT)
// rt[k] = exp(-2 pi i k/n) for k < n/2 (R2C/C2R split)
template <int N>
__device__ __forceinline__ void load_tw(float2* dst, const float2* __restrict__ src) {
  for (int i = threadIdx.x; i < N; i += blockDim.x) dst[i] = src[i];
}

// --------------------------------------------------------------- R2C / C2R split (in place on one row)
// Row holds M+1 complex slots. Forward: Z[0..M-1] -> X[0..M]. Inverse: X[0..M] -> Z[0..M-1]
// such that the unnormalised inverse M-point FFT of Z yields x[2m] + i x[2m+1]
// of the unnormalised inverse n-point DFT.
__device__ __forceinline__ void r2c_split(float2* row, int k, int M, const float2* rt) {
  if (k == 0) {
    float2 z = row[0];
    row[0] = make_float2(z.x + z.y, 0.f);
    row[M] = make_float2(z.x - z.y, 0.f);
    return;
  }
  float2 zk = row[k], zm = row[M - k];
  // E = (Zk + conj Zm)/2 ; O = (Zk - conj Zm)/(2i)
  float2 E = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
  float2 O = make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
  float2 w = rt[k];
  float2 WO = cmul(w, O);
  row[k] = make_float2(E.x + WO.x, E.y + WO.y);
  if (k != M - k) row[M - k] = make_float2(E.x - WO.x, -(E.y - WO.y));
}

__device__ __forceinline__ void c2r_split(float2* row, int k, int M, const float2* rt) {
  if (k == 0) {
    float a = row[0].x, b = row[M].x;
    row[0] = make_float2(a + b, a - b);
    return;
  }
  float2 xk = row[k], xm = row[M - k];
  // E' = Xk + conj Xm ; O' = (Xk - conj Xm) * conj(W^k) ; Z = E' + i O'
  float2 E = make_float2(xk.x + xm.x, xk.y - xm.y);
  float2 D = make_float2(xk.x - xm.x, xk.y + xm.y);
  float2 w = rt[k];
  float2 O = cmul(D, make_float2(w.x, -w.y));
  row[k] = make_float2(E.x - O.y, E.y + O.x);
  if (k != M - k) row[M - k] = make_float2(E.x + O.y, -E.y + O.x);
}

// ====================================================================== z passes
enum ZOp { Z_COPY = 0, Z_SQUARE = 1, Z_PREC_LOCAL = 2 };

// Shared memory of a z-pass tile: tw[M] | rt[M] | F*L rows of M+1 complex.
__device__ __forceinline__ float2* zrow(float2* tile, int f, int l, const Geo& g) {
  return tile + ((long)f * g.L + l) * (g.M + 1);
}

template <int M>
__device__ __forceinline__ void z_fft(float2* tile, int F, const Geo& g, const float2* tw, bool fwd) {
  Lines Ls{g.L, g.L, M + 1, 0, 1};
  for (int f = 0; f < F; ++f) {
    if (fwd) fft_lines<float2, M, true>(tile + (long)f * g.L * (M + 1), Ls, tw);
    else fft_lines<float2, M, false>(tile + (long)f * g.L * (M + 1), Ls, tw);
  }
}

__device__ __forceinline__ void z_split_all(float2* tile, int F, const Geo& g, const float2* rt, bool fwd) {
  const int half = g.M / 2 + 1;
  for (int e = threadIdx.x; e < F * g.L * half; e += blockDim.x) {
    const int k = e % half, fl = e / half;
    float2* row = tile + (long)fl * (g.M + 1);
    if (fwd) r2c_split(row, k, g.M, rt);
    else c2r_split(row, k, g.M, rt);
  }
  __syncthreads();
}

__device__ __forceinline__ void z_store_k(const float2* tile, float2* __restrict__ kout, int F, const Geo& g,
                                          long r0) {
  const int w = g.M + 1;
  for (int e = threadIdx.x; e < F * g.L * w; e += blockDim.x) {
    const int k = e % w, fl = e / w, l = fl % g.L, f = fl / g.L;
    kout[f * g.kfs + (r0 + l) * g.Pz + k] = tile[(long)fl * w + k];
  }
}

__device__ __forceinline__ void z_load_k(float2* tile, const float2* __restrict__ kin, int F, const Geo& g, long r0) {
  const int w = g.M + 1;
  for (int e = threadIdx.x; e < F * g.L * w; e += blockDim.x) {
    const int k = e % w, fl = e / w, l = fl % g.L, f = fl / g.L;
    tile[(long)fl * w + k] = kin[f * g.kfs + (r0 + l) * g.Pz + k];
  }
}

// Forward z pass: real fields -> pointwise op -> R2C -> half-spectrum rows.
// Martensite (Z_SQUARE): fields eta_p -> eta_p^2.
// Precipitate (Z_PREC_LOCAL): fields (dc, eta_1..V) -> (mu_c, F_1..F_V) with noise on eta.
template <int M, int OP>
__global__ void __launch_bounds__(1024) k_zfwd(const float* __restrict__ in, float2* __restrict__ kout, Geo g,
                                              int F, MartK mk, PrecK pk, const float2* __restrict__ twM,
                                              const float2* __restrict__ rtn) {
  extern __shared__ float2 smem[];
  float2* tw = smem;
  float2* rt = smem + M;
  float2* tile = smem + 2 * M;
  const long r0 = (long)blockIdx.x * g.L;
  load_tw<M>(tw, twM);
  load_tw<M>(rt, rtn);
  const int n = g.n;
  if constexpr (OP == Z_PREC_LOCAL) {
    // each thread handles 4 consecutive voxels (2 complex slots) of one row, all fields
    const int q4 = M / 2;
    for (int e = threadIdx.x; e < g.L * q4; e += blockDim.x) {
      const int l = e / q4, q = e % q4;
      const long vox = (r0 + l) * n + 4 * q;
      float v[8][4];
      for (int f = 0; f < F; ++f) {
        float4 t = *reinterpret_cast<const float4*>(in + f * g.nvox + vox);
        v[f][0] = t.x; v[f][1] = t.y; v[f][2] = t.z; v[f][3] = t.w;
      }
      const int V = pk.V;
      float xi[4] = {0.f, 0.f, 0.f, 0.f};
      float out[8][4];
      for (int i = 0; i < 4; ++i) {
        float c = pk.cbar + v[0][i];
        float s2 = 0.f;
        for (int a = 1; a <= V; ++a) s2 += v[a][i] * v[a][i];
        out[0][i] = 0.5f * pk.A2 * s2 + pk.A1 * (c - pk.c1);
      }
      for (int a = 1; a <= V; ++a) {
        if (pk.noise != 0.f) normal4(pk.seed, (uint32_t)(vox >> 2), (uint32_t)(a - 1), pk.step, xi);
        for (int i = 0; i < 4; ++i) {
          const float et = v[a][i];
          float t2 = 0.f, t4 = 0.f, t22 = 0.f;
          for (int b = 1; b <= V; ++b) {
            if (b == a) continue;
            const float qq = v[b][i] * v[b][i];
            t2 += qq; t4 += qq * qq;
          }
          // sum over ordered pairs b != d, both != a: (sum q)^2 - sum q^2
          t22 = t2 * t2 - t4;
          const float c = pk.cbar + v[0][i];
          const float e2 = et * et;
          out[a][i] = (2.f * pk.A5 * et + 4.f * pk.A6 * e2 * et) * t2 + 2.f * pk.A6 * et * t4 +
                      2.f * pk.A7 * et * t22 + pk.A2 * et * (c - pk.c2) - pk.A3 * e2 * et +
                      pk.A4 * e2 * e2 * et + pk.noise * xi[i];
        }
      }
      for (int f = 0; f < F; ++f) {
        float2* row = zrow(tile, f, l, g);
        row[2 * q] = make_float2(out[f][0], out[f][1]);
        row[2 * q + 1] = make_float2(out[f][2], out[f][3]);
      }
    }
  } else {
    for (int e = threadIdx.x; e < F * g.L * M; e += blockDim.x) {
      const int m = e % M, fl = e / M, l = fl % g.L, f = fl / g.L;
      float2 t = *reinterpret_cast<const float2*>(in + f * g.nvox + (r0 + l) * n + 2 * m);
      if (OP == Z_SQUARE) t = make_float2(t.x * t.x, t.y * t.y);
      tile[(long)fl * (M + 1) + m] = t;
    }
  }
  __syncthreads();
  z_fft<M>(tile, F, g, tw, true);
  z_split_all(tile, F, g, rt, true);
  z_store_k(tile, kout, F, g, r0);
}

// Martensite middle z pass: R_hat rows -> C2R -> F = f'(eta) + 2 w eta R + a xi -> R2C.
template <int M>
__global__ void __launch_bounds__(1024) k_zmid_mart(float2* __restrict__ kbuf, const float* __restrict__ eta, Geo g,
                                                   MartK mk, const float2* __restrict__ twM,
                                                   const float2* __restrict__ rtn) {
  extern __shared__ float2 smem[];
  float2* tw = smem;
  float2* rt = smem + M;
  float2* tile = smem + 2 * M;
  const long r0 = (long)blockIdx.x * g.L;
  const int F = mk.F, n = g.n;
  load_tw<M>(tw, twM);
  load_tw<M>(rt, rtn);
  z_load_k(tile, kbuf, F, g, r0);
  __syncthreads();
  z_split_all(tile, F, g, rt, false);
  z_fft<M>(tile, F, g, tw, false);
  const int q4 = M / 2;
  for (int e = threadIdx.x; e < g.L * q4; e += blockDim.x) {
    const int l = e / q4, q = e % q4;
    const long vox = (r0 + l) * n + 4 * q;
    float et[8][4];
    for (int f = 0; f < F; ++f) {
      float4 t = *reinterpret_cast<const float4*>(eta + f * g.nvox + vox);
      et[f][0] = t.x; et[f][1] = t.y; et[f][2] = t.z; et[f][3] = t.w;
    }
    float s2[4];
    for (int i = 0; i < 4; ++i) {
      float s = 0.f;
      for (int f = 0; f < F; ++f) s += et[f][i] * et[f][i];
      s2[i] = s;
    }
    for (int f = 0; f < F; ++f) {
      float2* row = zrow(tile, f, l, g);
      float2 ra = row[2 * q], rb = row[2 * q + 1];
      float R[4] = {ra.x, ra.y, rb.x, rb.y};
      float xi[4] = {0.f, 0.f, 0.f, 0.f};
      if (mk.noise != 0.f) normal4(mk.seed, (uint32_t)(vox >> 2), (uint32_t)f, mk.step, xi);
      float o[4];
      for (int i = 0; i < 4; ++i) {
        const float e1 = et[f][i];
        o[i] = mk.A * e1 - mk.B * e1 * e1 + mk.C * e1 * s2[i] + mk.two_wel * e1 * R[i] + mk.noise * xi[i];
      }
      row[2 * q] = make_float2(o[0], o[1]);
      row[2 * q + 1] = make_float2(o[2], o[3]);
    }
  }
  __syncthreads();
  z_fft<M>(tile, F, g, tw, true);
  z_split_all(tile, F, g, rt, true);
  z_store_k(tile, kbuf, F, g, r0);
}

// Inverse z pass: half-spectrum rows -> C2R -> real fields.
template <int M>
__global__ void __launch_bounds__(1024) k_zinv(const float2* __restrict__ kin, float* __restrict__ out, Geo g, int F,
                                              const float2* __restrict__ twM, const float2* __restrict__ rtn) {
  extern __shared__ float2 smem[];
  float2* tw = smem;
  float2* rt = smem + M;
  float2* tile = smem + 2 * M;
  const long r0 = (long)blockIdx.x * g.L;
  load_tw<M>(tw, twM);
  load_tw<M>(rt, rtn);
  z_load_k(tile, kin, F, g, r0);
  __syncthreads();
  z_split_all(tile, F, g, rt, false);
  z_fft<M>(tile, F, g, tw, false);
  for (int e = threadIdx.x; e < F * g.L * M; e += blockDim.x) {
    const int m = e % M, fl = e / M, l = fl % g.L, f = fl / g.L;
    *reinterpret_cast<float2*>(out + f * g.nvox + (r0 + l) * g.n + 2 * m) = tile[(long)fl * (M + 1) + m];
  }
}

}  // namespace gsk
