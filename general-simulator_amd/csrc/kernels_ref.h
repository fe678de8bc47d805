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

// The update of one voxel and variant: e = eta_p(x), xp..zm its six neighbours, s2 = sum_i
// (sc eta_i)^2 and s1 = sum_i eta_i over the variants at x, r = R_p(x), kv = K_p(x), xi the
// normal deviate. The four force terms w_grad Grad, Chem, El, Dis go to f[0..3].
__device__ __forceinline__ float ref_mart_one(const RefMartK& k, float e, float xp, float xm, float yp, float ym,
                                              float zp, float zm, float s2, float s1, float r, float kv, float xi,
                                              float (&f)[4]) {
  const float grad = (xp + xm - 2.f * e) * k.gx + (yp + ym - 2.f * e) * k.gy + (zp + zm - 2.f * e) * k.gz;
  float chem;
  if (k.chem_form == 0)
    chem = e >= 0.f ? -k.wC * (k.sc * e) * (k.a1 - k.a2 * (k.sc * e) * (k.sc * e) + k.a3 * s2)
                    : 1000.f * k.wC * k.a2 * e * e;
  else
    chem = -k.wC * e * (k.a1 - k.a2 * e + k.a3 * s1);
  const float el = -k.wE * r * e;
  const float dis = -k.wD * kv * e;
  f[0] = k.wG * grad;
  f[1] = chem;
  f[2] = el;
  f[3] = dis;
  return e + k.dt * (k.wG * grad + chem + el + dis + k.wN * xi);
}
__device__ __forceinline__ void ref_mart_acc(double (&fs)[4], const float (&f)[4]) {
  for (int j = 0; j < 4; ++j) fs[j] += (double)f[j];
}

// One thread per 4 consecutive z voxels of an nx x ny x nz grid (nz a multiple of 4), any
// number of variants; eta: V fields, RE: V fields holding R on entry and the new eta on exit
// (each element is read before it is overwritten by the same thread).
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
        float f[4];
        float nv = ref_mart_one(k, ev[i], ev[i + dxp], ev[i + dxm], ev[i + dyp], ev[i + dym], ev[zp], ev[zm], S2[j],
                                S1[j], RE[v * N + i], K ? K[v * N + i] : 0.f, xi[j], f);
        if (fpart) ref_mart_acc(fs, f);
        if (defect) nv *= 1.f - defect[i];
        RE[v * N + i] = nv;
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

// 16-byte non-temporal load and store (data read or written once, kept out of the caches' way)
typedef float nt_f4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ld_nt4(const float4* p) {
  const nt_f4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_f4*>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void st_nt4(float4* p, float4 v) {
  const nt_f4 w = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(w, reinterpret_cast<nt_f4*>(p));
}

// The same update for V = 1..4 variants with 16-byte loads and stores: a chunk is 256
// consecutive quads (1024 voxels) per work group. The chunks are split into eight contiguous
// ranges, one per XCD (work group index mod 8, the order in which work groups are dealt to the
// XCDs). Chunks are numbered in-plane position first, plane second (when a plane holds whole
// chunks), so the work groups of an XCD at one time hold the same rows of consecutive planes: the
// x neighbours of a chunk are read by the groups beside it, the y neighbours by the previous and
// next sweep, out of that XCD's L2. R, K, the defect mask and the new eta are read or written
// once: non-temporal accesses keep them from evicting the eta that is read again as neighbours.
// ZS: the z rows are L = nz/4 lanes of one wave (L divides 64) and every chunk is whole: the z
// neighbours across a quad boundary come from the neighbouring lane (a lane shuffle, periodic in
// the row) instead of a load. Work groups: a multiple of 8, each writes fpart[blockIdx.x*4 + j].
template <int V, bool ZS>
__global__ void __launch_bounds__(256) k_ref_mart_q(const float* __restrict__ eta, float* __restrict__ RE,
                                                    const float* __restrict__ defect, const float* __restrict__ K,
                                                    int nx, int ny, int nz, RefMartK k, double* __restrict__ fpart) {
  double fs[4] = {0.0, 0.0, 0.0, 0.0};
  const unsigned nq = (unsigned)nx * (unsigned)ny * (unsigned)nz / 4u, L = (unsigned)nz / 4u;
  const unsigned qp = (unsigned)ny * L;  // quads per x-plane
  const unsigned C = (nq + 255u) / 256u, G8 = gridDim.x / 8u, xcd = blockIdx.x & 7u;
  const unsigned c1 = (unsigned)((unsigned long long)C * (xcd + 1u) / 8u);
  const uint32_t step = step_of(k.step, k.stepp);
  const float4* e4 = reinterpret_cast<const float4*>(eta);
  float4* r4 = reinterpret_cast<float4*>(RE);
  const unsigned lane = threadIdx.x & 63u, lr = lane & (L - 1u), rowb = lane - lr;
  // chunk order: the CP chunks of a plane (CP = qp / 256 when whole) are walked plane by plane for
  // each in-plane chunk, so that consecutive chunks are x-neighbours
  const unsigned CP = qp % 256u == 0u ? qp / 256u : 0u;
  for (unsigned c = (unsigned)((unsigned long long)C * xcd / 8u) + blockIdx.x / 8u; c < c1; c += G8) {
    const unsigned ch = CP ? (c % (unsigned)nx) * CP + c / (unsigned)nx : c;
    const unsigned qd = ch * 256u + threadIdx.x;
    if (!ZS && qd >= nq) continue;
    const unsigned x = qd / qp, rem = qd - x * qp, y = rem / L, zq = rem - y * L;
    const unsigned qxp = x + 1u == (unsigned)nx ? qd - x * qp : qd + qp;
    const unsigned qxm = x == 0u ? qd + ((unsigned)nx - 1u) * qp : qd - qp;
    const unsigned qyp = y + 1u == (unsigned)ny ? qd - y * L : qd + L;
    const unsigned qym = y == 0u ? qd + ((unsigned)ny - 1u) * L : qd - L;
    float4 E[V];
    float S2[4] = {0.f, 0.f, 0.f, 0.f}, S1[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int u = 0; u < V; ++u) {
      E[u] = e4[u * nq + qd];
      const float ev[4] = {E[u].x, E[u].y, E[u].z, E[u].w};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        S2[j] += (k.sc * ev[j]) * (k.sc * ev[j]);
        S1[j] += ev[j];
      }
    }
    const float4 D =
        defect ? ld_nt4(reinterpret_cast<const float4*>(defect) + qd) : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const unsigned o = (unsigned)v * nq;
      const float4 xp = e4[o + qxp], xm = e4[o + qxm], yp = e4[o + qyp], ym = e4[o + qym];
      const float4 R = ld_nt4(r4 + o + qd);
      const float4 Kv =
          K ? ld_nt4(reinterpret_cast<const float4*>(K) + o + qd) : make_float4(0.f, 0.f, 0.f, 0.f);
      float zlo, zhi;  // eta at z0 - 1 and z0 + 4 (periodic)
      if (ZS) {
        zlo = __shfl(E[v].w, (int)(rowb + ((lr - 1u) & (L - 1u))), 64);
        zhi = __shfl(E[v].x, (int)(rowb + ((lr + 1u) & (L - 1u))), 64);
      } else {
        const unsigned i0 = 4u * qd, rs = i0 - 4u * zq, z0 = 4u * zq;  // rs: start of the z row
        zlo = eta[(size_t)o * 4u + rs + (z0 == 0u ? (unsigned)nz - 1u : z0 - 1u)];
        zhi = eta[(size_t)o * 4u + rs + (z0 + 4u == (unsigned)nz ? 0u : z0 + 4u)];
      }
      float xi[4] = {0.f, 0.f, 0.f, 0.f};
      if (k.wN != 0.f) normal4(k.seed, qd, (uint32_t)v, step, xi);
      const float4 e = E[v];
      float4 nv;
      float f0[4], f1[4], f2[4], f3[4];
      nv.x = ref_mart_one(k, e.x, xp.x, xm.x, yp.x, ym.x, e.y, zlo, S2[0], S1[0], R.x, Kv.x, xi[0], f0);
      nv.y = ref_mart_one(k, e.y, xp.y, xm.y, yp.y, ym.y, e.z, e.x, S2[1], S1[1], R.y, Kv.y, xi[1], f1);
      nv.z = ref_mart_one(k, e.z, xp.z, xm.z, yp.z, ym.z, e.w, e.y, S2[2], S1[2], R.z, Kv.z, xi[2], f2);
      nv.w = ref_mart_one(k, e.w, xp.w, xm.w, yp.w, ym.w, zhi, e.z, S2[3], S1[3], R.w, Kv.w, xi[3], f3);
      if (fpart) {
        ref_mart_acc(fs, f0);
        ref_mart_acc(fs, f1);
        ref_mart_acc(fs, f2);
        ref_mart_acc(fs, f3);
      }
      if (defect) {
        nv.x *= 1.f - D.x;
        nv.y *= 1.f - D.y;
        nv.z *= 1.f - D.z;
        nv.w *= 1.f - D.w;
      }
      st_nt4(r4 + o + qd, nv);
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
