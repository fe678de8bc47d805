// Plane-resident middle pass of the FP32 fused martensite step, edges n = 32, 64, 128.
//
// One work group holds one x-plane of one variant's half spectrum in LDS: n rows (ky) of
// n/2 + 1 values (kz < n/2, then the Nyquist value), at an odd row pitch. On it the group runs
// what the y inverse -> middle z pass -> y forward chain does through HBM:
//   y inverse of R_p on the n/2 + 1 columns;
//   C2R split and M-point inverse along each row (M = n/2), giving R_p in real space;
//   F_p = A eta_p - B eta_p^2 + C eta_p sum_q eta_q^2 + 2 w eta_p R_p + a xi_p on every voxel,
//     eta of all variants read as 16-byte quads, xi from the same Philox4x32-10 blocks;
//   M-point forward along each row and R2C split; y forward; F_p written back in place.
// HBM per voxel and variant: the R_p and F_p spectra and the V real eta fields, 8 + 4V bytes
// against 28 for the three kernels. At n = 128 a plane takes 66.5 KB of LDS (two groups per CU).
// The passes are the engine's LDS Stockham passes (fft_core.h) with the y and z plans of the
// strided and contiguous kernels.
#pragma once
#include "kernels_z.h"

namespace gsk {

template <int N> constexpr bool plane_mid_size() { return N == 32 || N == 64 || N == 128; }

// Row pitch of the plane (complex elements): odd and >= n/2 + 1. Column passes put consecutive
// lanes on consecutive columns (unit stride); row passes put them on consecutive rows, which an
// odd pitch spreads over distinct bank pairs.
__host__ __device__ constexpr int plane_pitch(int N) { return (N / 2 + 1) | 1; }

// Threads: one per butterfly of the smallest radix, the larger of the y passes over the n/2 main
// columns and the z passes over the n rows.
template <int N> constexpr int plane_threads() {
  constexpr int M = N / 2;
  constexpr int a = M * (N / PlanOf<N>::rmin), b = N * (M / ZPlanOf<M>::rmin);
  constexpr int t = a > b ? a : b;
  return t < 64 ? 64 : (t + 63) / 64 * 64;
}
template <int N> constexpr size_t plane_lds() {
  return (size_t)(N + 2 * (N / 2) + N * plane_pitch(N)) * sizeof(float2);
}

// Grid: (n planes, V variants[, members]) through geo_tile / geo_field. VT: the variant count
// (1..4). mt (may be null): batched ensemble launch, member blockIdx.z's buffers and constants.
template <int N, int VT>
__global__ void __launch_bounds__(plane_threads<N>()) k_plane_mid(float2* __restrict__ kbuf, const float* __restrict__ eta,
                                                              Geo g, MartK mk, const float2* __restrict__ twN,
                                                              const float2* __restrict__ twM,
                                                              const float2* __restrict__ rtn,
                                                              const MemberPtrs* __restrict__ mt) {
  if (mt) {
    kbuf = mt[blockIdx.z].kbuf;
    eta = mt[blockIdx.z].fld;
    mk = mt[blockIdx.z].mk;
  }
  static_assert(plane_mid_size<N>() && VT >= 1 && VT <= 4, "plane pass: n = 32, 64, 128 and 1..4 variants");
  constexpr int M = N / 2, P = plane_pitch(N), Q = M / 2;
  constexpr int T = plane_threads<N>();
  extern __shared__ float2 smem[];
  float2* tw = smem;           // N-point table (y)
  float2* twz = smem + N;      // M-point table (z)
  float2* rt = smem + N + M;   // R2C / C2R split factors
  float2* pl = smem + N + 2 * M;
  const int p = geo_field(g);
  const unsigned rb = (unsigned)geo_tile(g) * N;  // row index (x*n + y) of the plane's ky = 0
  const int tid = threadIdx.x;
  float2* field = kbuf + p * g.kfs;
  for (int e = tid; e < N * M; e += T) pl[(e / M) * P + e % M] = at32(field, rb * M + (unsigned)e);
  for (int y = tid; y < N; y += T) pl[y * P + M] = at32(field, (unsigned)g.nq + rb + y);
  for (int i = tid; i < N; i += T) tw[i] = twN[i];
  for (int i = tid; i < M; i += T) {
    twz[i] = twM[i];
    rt[i] = rtn[i];
  }
  __syncthreads();

  const Lines cols{M, M, 1, 0, P}, nyq{1, 1, 0, 0, P}, rows{N, N, P, 0, 1};
  fft_lines_p<float2, N, false, typename PlanOf<N>::type>(pl, cols, tw);
  fft_lines_p<float2, N, false, typename PlanOf<N>::type>(pl + M, nyq, tw);

  // C2R split of every row (task (y, k) owns the pair k, M - k; k = 0 also owns M/2 and M)
  for (int e = tid; e < N * Q; e += T) {
    float2* z = pl + (e / Q) * P;
    const int k = e % Q;
    if (k == 0) {
      const float2 a = z[0], b = z[M], c = z[Q];
      float2 zh, unused;
      c2r_pair(c, c, rt[Q], zh, unused);
      z[0] = make_float2(a.x + b.x, a.x - b.x);
      z[Q] = zh;
    } else {
      float2 zk, zm;
      c2r_pair(z[k], z[M - k], rt[k], zk, zm);
      z[k] = zk;
      z[M - k] = zm;
    }
  }
  __syncthreads();
  fft_lines_p<float2, M, false, typename ZPlanOf<M>::type>(pl, rows, twz);

  // forces on quads of voxels: packed element 2q, 2q+1 of row y holds R at voxels 4q..4q+3
  const uint32_t step = step_of(mk.step, mk.stepp);
  const float* eb = eta + (size_t)rb * N;
  for (int e = tid; e < N * Q; e += T) {
    const int y = e / Q, q = e % Q;
    float2* z = pl + y * P;
    const unsigned vo = (unsigned)(y * N + 4 * q);  // voxel offset inside the plane
    float4 ep = make_float4(0.f, 0.f, 0.f, 0.f), s2 = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int f = 0; f < VT; ++f) {
      const float4 t = at32(reinterpret_cast<const float4*>(eb + f * g.nvox), vo / 4);
      s2.x += t.x * t.x; s2.y += t.y * t.y; s2.z += t.z * t.z; s2.w += t.w * t.w;
      if (f == p) ep = t;
    }
    float xi[4] = {0.f, 0.f, 0.f, 0.f};
    if (mk.noise != 0.f) normal4(mk.seed, (rb * N + vo) >> 2, (uint32_t)p, step, xi);
    const float2 ra = z[2 * q], rbv = z[2 * q + 1];
    const float e4[4] = {ep.x, ep.y, ep.z, ep.w}, s4[4] = {s2.x, s2.y, s2.z, s2.w}, r4[4] = {ra.x, ra.y, rbv.x, rbv.y};
    float o[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float et = e4[i];
      o[i] = mk.A * et - mk.B * et * et + mk.C * et * s4[i] + mk.two_wel * et * r4[i] + mk.noise * xi[i];
    }
    z[2 * q] = make_float2(o[0], o[1]);
    z[2 * q + 1] = make_float2(o[2], o[3]);
  }
  __syncthreads();
  fft_lines_p<float2, M, true, typename ZPlanOf<M>::type>(pl, rows, twz);

  // R2C split of every row back to n/2 + 1 half-spectrum values
  for (int e = tid; e < N * Q; e += T) {
    float2* z = pl + (e / Q) * P;
    const int k = e % Q;
    if (k == 0) {
      const float2 z0 = z[0], zh = z[Q];
      z[0] = make_float2(z0.x + z0.y, 0.f);
      z[M] = make_float2(z0.x - z0.y, 0.f);
      z[Q] = make_float2(zh.x, -zh.y);
    } else {
      float2 xk, xm;
      r2c_pair(z[k], z[M - k], rt[k], xk, xm);
      z[k] = xk;
      z[M - k] = xm;
    }
  }
  __syncthreads();
  fft_lines_p<float2, N, true, typename PlanOf<N>::type>(pl, cols, tw);
  fft_lines_p<float2, N, true, typename PlanOf<N>::type>(pl + M, nyq, tw);

  for (int e = tid; e < N * M; e += T) at32(field, rb * M + (unsigned)e) = pl[(e / M) * P + e % M];
  for (int y = tid; y < N; y += T) at32(field, (unsigned)g.nq + rb + y) = pl[y * P + M];
}

}  // namespace gsk
