// Reference rule of the martensite plugin (src/dynamics_mart.cu:181-417), FP32 like the reference:
//   eta += dt (w_grad Grad + Chem + El + Dis + w_noise xi), then eta *= (1 - defect)
//   Grad = sum_d (eta(x+e_d) + eta(x-e_d) - 2 eta(x)) / (2 Delta_d) / 3     (:181-188, periodic)
//   Chem = HEAD side: s = sqrt|a1/(a2-a3)|; eta >= 0: -w_chem (s eta)(a1 - a2 (s eta)^2 + a3 sum_i (s eta_i)^2),
//          eta < 0: 1000 w_chem a2 eta^2;  master side: -w_chem eta (a1 - a2 eta + a3 sum_i eta_i)   (:222-250)
//   El   = -w_el R_p eta_p with R_p = IFFT[sum_q B_pq FFT(eta_q^2)]/N over all q (HEAD, :270-289)
//   Dis  = -w_dis K_p eta_p (K_p = 25 sum_ij sigma_ij eps^p_ij, set once, :305-331)
// The elastic R comes from the fused transforms (zfwd square, y, x contraction, y, z inverse).
#pragma once
#include "kernels_aux.h"

namespace gsk {

struct RefMartK {
  float a1, a2, a3, sc;  // sc = sqrt|a1/(a2-a3)|
  float wG, wC, wE, wD, wN, dt;
  float gx, gy, gz;  // 1/(2 Delta_d)/3 per axis
  int V, chem_form;
  uint32_t step;
  const uint32_t* stepp;
  unsigned long long seed;
};

// One thread per 4 consecutive z voxels of an nx x ny x nz grid (nz a multiple of 4); eta: V
// fields, RE: V fields holding R on entry and
// the new eta on exit (each element is read before it is overwritten by the same thread).
// fpart (may be null): per-block FP64 sums of the four force terms of this step, w_grad Grad,
// Chem, El, Dis, over every voxel and variant (the reference's GradientForce, ChemicalForce,
// ElasticForce, DislocationForce arrays that Get averages, src/dynamics_mart.cu:443-456).
template <int = 0>
__global__ void k_ref_mart(const float* __restrict__ eta, float* __restrict__ RE, const float* __restrict__ defect,
                           const float* __restrict__ K, int nx, int ny, int nz, RefMartK k,
                           double* __restrict__ fpart) {
  double fs[4] = {0.0, 0.0, 0.0, 0.0};
  const long N = (long)nx * ny * nz, nq = N / 4;
  const long nn = (long)ny * nz;
  const uint32_t step = step_of(k.step, k.stepp);
  for (long qd = blockIdx.x * (long)blockDim.x + threadIdx.x; qd < nq; qd += (long)gridDim.x * blockDim.x) {
    const long i0 = qd * 4;
    const int x = (int)(i0 / nn), y = (int)((i0 / nz) % ny), z0 = (int)(i0 % nz);
    const long dxp = ((x + 1) % nx - x) * nn, dxm = ((x + nx - 1) % nx - x) * nn;
    const long dyp = (long)((y + 1) % ny - y) * nz, dym = (long)((y + ny - 1) % ny - y) * nz;
    float S2[4] = {0.f, 0.f, 0.f, 0.f}, S1[4] = {0.f, 0.f, 0.f, 0.f};
    for (int u = 0; u < k.V; ++u)
      for (int j = 0; j < 4; ++j) {
        const float e = eta[u * N + i0 + j];
        S2[j] += (k.sc * e) * (k.sc * e);
        S1[j] += e;
      }
    for (int v = 0; v < k.V; ++v) {
      const float* ev = eta + v * N;
      float xi[4] = {0.f, 0.f, 0.f, 0.f};
      if (k.wN != 0.f) normal4(k.seed, (uint32_t)qd, (uint32_t)v, step, xi);
      for (int j = 0; j < 4; ++j) {
        const long i = i0 + j;
        const int z = z0 + j;
        const long zp = i + ((z + 1) % nz - z), zm = i + ((z + nz - 1) % nz - z);
        const float e = ev[i];
        const float grad = (ev[i + dxp] + ev[i + dxm] - 2.f * e) * k.gx + (ev[i + dyp] + ev[i + dym] - 2.f * e) * k.gy +
                           (ev[zp] + ev[zm] - 2.f * e) * k.gz;
        float chem;
        if (k.chem_form == 0)
          chem = e >= 0.f ? -k.wC * (k.sc * e) * (k.a1 - k.a2 * (k.sc * e) * (k.sc * e) + k.a3 * S2[j])
                          : 1000.f * k.wC * k.a2 * e * e;
        else
          chem = -k.wC * e * (k.a1 - k.a2 * e + k.a3 * S1[j]);
        const float el = -k.wE * RE[v * N + i] * e;
        const float dis = K ? -k.wD * K[v * N + i] * e : 0.f;
        float nv = e + k.dt * (k.wG * grad + chem + el + dis + k.wN * xi[j]);
        if (defect) nv *= 1.f - defect[i];
        RE[v * N + i] = nv;
        if (fpart) {
          fs[0] += (double)(k.wG * grad);
          fs[1] += (double)chem;
          fs[2] += (double)el;
          fs[3] += (double)dis;
        }
      }
    }
  }
  if (fpart) {
    __shared__ double sh[16];
    for (int j = 0; j < 4; ++j) {
      const double r = block_sum(fs[j], sh);
      if (threadIdx.x == 0) fpart[blockIdx.x * 4 + j] = r;
    }
  }
}

// dst[i] = src[i] for count floats (16-byte vectors; count is a multiple of 4)
template <int = 0>
__global__ void k_copy4(float* __restrict__ dst, const float* __restrict__ src, long count) {
  const long n4 = count / 4;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n4; i += (long)gridDim.x * blockDim.x)
    reinterpret_cast<float4*>(dst)[i] = reinterpret_cast<const float4*>(src)[i];
}

// advances the device step counter (the reference rule's last launch of a step)
template <int = 0>
__global__ void k_advance(uint32_t* __restrict__ stepp) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *stepp += 1u;
}

}  // namespace gsk
