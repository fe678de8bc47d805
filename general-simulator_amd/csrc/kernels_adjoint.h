// Adjoint (reverse-mode) kernels of the FP32 martensite step.
//
// One forward step in real space is eta' = D * (eta - dt L F(eta)), D(k) = 1/(1 + dt L kappa k^2),
// F_p = A eta_p - B eta_p^2 + C eta_p S + 2 w eta_p R_p + a xi_p, R_p = B_pq * eta_q^2 (a
// convolution with the Khachaturyan tensor), S = sum_q eta_q^2. For lam = dLoss/deta':
//   mu     = D * lam
//   lam_p <- mu_p - dt L [ (A - 2B eta_p + C S + 2w R_p) mu_p + 2C eta_p sum_q eta_q mu_q + 4w eta_p Q_p ],
//   Q_p    = B_pr * (eta_r mu_r)
// and the parameter gradients collect k-space inner products (host side in FP64):
//   dLoss/dL     += -dt kappa X - dt <mu, F>,  dLoss/dkappa += -dt L X,
//   X = (1/N) sum_k Re(conj(lam_hat) eta'_hat) k^2 D
//   dLoss/ds_p   += -2 dt L w sum_q (G_pq + G_qp),  G_rq = (1/N) sum_k Re(conj(P_hat_r) B_rq S_hat_q)
// The transforms are the library's FP32 R2C/C2R passes on natural-layout spectra
// [f][kx][ky][kz], kz <= n/2 (index kidx); sums over the full spectrum use the
// half-spectrum weights w(kz) = 1 for kz in {0, n/2}, 2 otherwise.
#pragma once
#include "kernels_aux.h"

namespace gsk {

struct AdjK {
  float A, B, C, two_wel, noise, dtL, dtLk, kfac, inv_n3;
  int V;
  uint32_t step;
  unsigned long long seed;
};

__device__ __forceinline__ void half_index(long i, const Geo& g, int& kx, int& ky, int& kz, long& row) {
  const int H = g.M + 1;
  kz = (int)(i % H);
  row = i / H;
  kx = (int)(row / g.n);
  ky = (int)(row % g.n);
}

// Packed B_pq (p <= q) on the half spectrum in natural layout, FP32: Bh[kidx(g, pq, row, kz)].
template <int = 0>
__global__ void k_btensor_half(float* __restrict__ Bh, Geo g, BArgs a) {
  const long total = (long)g.n * g.n * (g.M + 1);
  const int V = a.V1;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    int kx, ky, kz;
    long row;
    half_index(i, g, kx, ky, kz, row);
    double out[64];
    btensor_half_spec(kx, ky, kz, g.n, a, out);
    for (int p = 0; p < V; ++p)
      for (int q = p; q < V; ++q) {
        const int pq = p * V - (p * (p - 1)) / 2 + (q - p);
        Bh[kidx(g, pq, row, kz)] = (float)out[p * V + q];
      }
  }
}

// lam_hat <- D lam_hat / N in place (fields 0..V-1); part[block] = sum w Re(conj(lam) eta1) k^2 D.
template <int = 0>
__global__ void k_adj_dmul(float2* __restrict__ lam, const float2* __restrict__ eta1, Geo g, AdjK k,
                           double* __restrict__ part) {
  __shared__ double sh[16];
  const long per = (long)g.n * g.n * (g.M + 1), total = per * k.V;
  double acc = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int f = (int)(i / per);
    int kx, ky, kz;
    long row;
    half_index(i % per, g, kx, ky, kz, row);
    const long idx = kidx(g, f, row, kz);
    const float kxv = kcomp(kx, g.n, k.kfac), kyv = kcomp(ky, g.n, k.kfac), kzv = k.kfac * (float)kz;
    const float k2 = kxv * kxv + kyv * kyv + kzv * kzv;
    const float D = 1.f / (1.f + k.dtLk * k2);
    const float2 l = lam[idx], e = eta1[idx];
    const double w = (kz == 0 || kz == g.M) ? 1.0 : 2.0;
    acc += w * ((double)l.x * e.x + (double)l.y * e.y) * (double)k2 * (double)D;
    lam[idx] = make_float2(l.x * D * k.inv_n3, l.y * D * k.inv_n3);
  }
  const double r = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// kb fields 0..V-1 hold S_hat (eta^2), fields V..2V-1 P_hat (eta mu). In place:
// fields 0..V-1 <- R_hat = B S_hat / N, fields V..2V-1 <- Q_hat = B P_hat / N.
// part[block*V*V + r*V + q] = sum w Re(conj(P_r) B_rq S_q).
template <int V>
__global__ void k_adj_contract(float2* __restrict__ kb, const float* __restrict__ Bh, Geo g, float inv_n3,
                               double* __restrict__ part) {
  __shared__ double sh[16];
  const long per = (long)g.n * g.n * (g.M + 1);
  double G[V * V];
#pragma unroll
  for (int i = 0; i < V * V; ++i) G[i] = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < per; i += (long)gridDim.x * blockDim.x) {
    int kx, ky, kz;
    long row;
    half_index(i, g, kx, ky, kz, row);
    const long off = kidx(g, 0, row, kz);
    const double w = (kz == 0 || kz == g.M) ? 1.0 : 2.0;
    float2 S[V], P[V];
    float b[V * (V + 1) / 2];
#pragma unroll
    for (int q = 0; q < V; ++q) {
      S[q] = kb[q * g.kfs + off];
      P[q] = kb[(V + q) * g.kfs + off];
    }
#pragma unroll
    for (int pq = 0; pq < V * (V + 1) / 2; ++pq) b[pq] = Bh[pq * g.kfs + off];
#pragma unroll
    for (int p = 0; p < V; ++p) {
      float2 R = make_float2(0.f, 0.f), Q = make_float2(0.f, 0.f);
#pragma unroll
      for (int q = 0; q < V; ++q) {
        const int lo = p < q ? p : q, hi = p < q ? q : p;
        const float bv = b[lo * V - (lo * (lo - 1)) / 2 + (hi - lo)];
        R.x += bv * S[q].x; R.y += bv * S[q].y;
        Q.x += bv * P[q].x; Q.y += bv * P[q].y;
        G[p * V + q] += w * (double)bv * ((double)P[p].x * S[q].x + (double)P[p].y * S[q].y);
      }
      kb[p * g.kfs + off] = make_float2(R.x * inv_n3, R.y * inv_n3);
      kb[(V + p) * g.kfs + off] = make_float2(Q.x * inv_n3, Q.y * inv_n3);
    }
  }
#pragma unroll
  for (int i = 0; i < V * V; ++i) {
    const double r = block_sum(G[i], sh);
    if (threadIdx.x == 0) part[blockIdx.x * V * V + i] = r;
  }
}

// out[q] = eta_q^2, out[V + q] = eta_q mu_q (real fields of N voxels)
template <int = 0>
__global__ void k_adj_prep(const float* __restrict__ eta, const float* __restrict__ mu, float* __restrict__ out, int V,
                           long N) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x)
    for (int q = 0; q < V; ++q) {
      const float e = eta[q * N + i];
      out[q * N + i] = e * e;
      out[(V + q) * N + i] = e * mu[q * N + i];
    }
}

// lam <- mu - dt L J^T mu at eta (RQ: R then Q, real); part[block] = <mu, F(eta)> including noise.
// One thread per 4 consecutive voxels (one Philox block of noise per variant).
template <int = 0>
__global__ void k_adj_lambda(const float* __restrict__ eta, const float* __restrict__ mu, const float* __restrict__ RQ,
                             float* __restrict__ lam, long N, AdjK k, double* __restrict__ part) {
  __shared__ double sh[16];
  const int V = k.V;
  double acc = 0.0;
  for (long qd = blockIdx.x * (long)blockDim.x + threadIdx.x; qd < N / 4; qd += (long)gridDim.x * blockDim.x) {
    const long v0 = qd * 4;
    float S[4] = {0.f, 0.f, 0.f, 0.f}, T[4] = {0.f, 0.f, 0.f, 0.f};
    for (int q = 0; q < V; ++q)
      for (int j = 0; j < 4; ++j) {
        const float e = eta[q * N + v0 + j];
        S[j] += e * e;
        T[j] += e * mu[q * N + v0 + j];
      }
    for (int p = 0; p < V; ++p) {
      float xi[4] = {0.f, 0.f, 0.f, 0.f};
      if (k.noise != 0.f) normal4(k.seed, (uint32_t)qd, (uint32_t)p, k.step, xi);
      for (int j = 0; j < 4; ++j) {
        const long i = p * N + v0 + j;
        const float e = eta[i], m = mu[i], R = RQ[i], Q = RQ[V * N + i];
        const float F = k.A * e - k.B * e * e + k.C * e * S[j] + k.two_wel * e * R + k.noise * xi[j];
        acc += (double)m * (double)F;
        const float jt = (k.A - 2.f * k.B * e + k.C * S[j] + k.two_wel * R) * m + 2.f * k.C * e * T[j] +
                         2.f * k.two_wel * e * Q;
        lam[i] = m - k.dtL * jt;
      }
    }
  }
  const double r = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// Full (not packed) rectangular B_pq, p < V1, q < V2, on the half spectrum in natural layout:
// Bh[kidx(g, p*V2 + q, row, kz)] (the stress plugin's 6 base strains x nvar variants).
template <int = 0>
__global__ void k_btensor_half_rect(float* __restrict__ Bh, Geo g, BArgs a) {
  const long total = (long)g.n * g.n * (g.M + 1);
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    int kx, ky, kz;
    long row;
    half_index(i, g, kx, ky, kz, row);
    double out[64];
    btensor_half_spec(kx, ky, kz, g.n, a, out);
    for (int pq = 0; pq < a.V1 * a.V2; ++pq) Bh[kidx(g, pq, row, kz)] = (float)out[pq];
  }
}

// In place on natural-layout spectra: fields 0..V2-1 hold X_q; fields 0..V1-1 receive
// scale * sum_q B_pq X_q (every input of a wave vector is read before any output is written).
template <int = 0>
__global__ void k_contract_rect(float2* __restrict__ kb, const float* __restrict__ Bh, Geo g, int V1, int V2,
                                float scale) {
  const long per = (long)g.n * g.n * (g.M + 1);
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < per; i += (long)gridDim.x * blockDim.x) {
    int kx, ky, kz;
    long row;
    half_index(i, g, kx, ky, kz, row);
    const long off = kidx(g, 0, row, kz);
    float2 X[8];
    for (int q = 0; q < V2; ++q) X[q] = kb[q * g.kfs + off];
    float2 Y[8];
    for (int p = 0; p < V1; ++p) {
      float2 acc = make_float2(0.f, 0.f);
      for (int q = 0; q < V2; ++q) {
        const float bv = Bh[(p * V2 + q) * g.kfs + off];
        acc.x += bv * X[q].x;
        acc.y += bv * X[q].y;
      }
      Y[p] = make_float2(acc.x * scale, acc.y * scale);
    }
    for (int p = 0; p < V1; ++p) kb[p * g.kfs + off] = Y[p];
  }
}

// x <- x^2 on count floats (the eta^2 source)
template <int = 0>
__global__ void k_square(float* __restrict__ x, long count) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < count; i += (long)gridDim.x * blockDim.x)
    x[i] = x[i] * x[i];
}

}  // namespace gsk
