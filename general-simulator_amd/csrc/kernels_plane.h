// Plane-resident passes of the FP32 fused martensite step, edges n = 32, 64, 128.
//
// One work group holds one x-plane of one field's half spectrum in LDS: n rows (ky) of
// n/2 + 1 values (kz < n/2, then the Nyquist value), at an odd row pitch, and runs on it what
// a y pass, a z pass and a y pass do through HBM. At n = 128 a plane takes 66.5 KB of LDS (two
// groups per CU). The passes are the engine's LDS Stockham passes (fft_core.h) with the plans
// of the strided (y) and contiguous (z) kernels.
//
// k_plane_mid (middle of a step, per variant p): y inverse of R_p, C2R split and M-point inverse
//   along each row (M = n/2), F_p = A eta_p - B eta_p^2 + C eta_p sum_q eta_q^2 + 2 w eta_p R_p
//   + a xi_p on every voxel (eta of all variants read as 16-byte quads, xi from the same
//   Philox4x32-10 blocks), M-point forward along each row and R2C split, y forward, F_p
//   written back in place. HBM per voxel and variant: 8 + 4V bytes against 28 for the chain.
// k_plane_sq (end of a step, per field): y inverse of eta_hat, C2R, eta written to the real
//   field, squared, R2C, y forward: the y,z-transformed eta^2 the next step starts from (the
//   chain leaves it z-transformed only). 12 bytes per voxel and field against 28.
// k_plane_sq0: the same y,z-transformed eta^2 from the real eta fields, bit-identical to what
//   k_plane_sq leaves (the squares are taken of the same FP32 values, then the same code).
#pragma once
#include "kernels_z.h"

namespace gsk {

template <int N> constexpr bool plane_size() { return N == 32 || N == 64 || N == 128; }

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
// One radix-R Stockham pass over the lines of L with up to K butterflies per thread of a
// T-thread group: the same butterflies and arithmetic (pass_math) as fft_pass, which takes one.
template <typename T2, int N, int R, bool FWD, int K, int T>
__device__ __forceinline__ void pass_k(T2* lds, const Lines& L, int Ns, const T2* tw) {
  const int nb = L.nl * (N / R);
  T2 v[K][R];
  int base[K], j[K];
#pragma unroll
  for (int t = 0; t < K; ++t) {
    const int b = threadIdx.x + t * T;
    j[t] = 0;
    base[t] = 0;
    if (b < nb) {
      j[t] = b / L.nl;
      base[t] = L.base(b % L.nl);
#pragma unroll
      for (int r = 0; r < R; ++r) v[t][r] = lds[base[t] + (j[t] + r * (N / R)) * L.ps];
    }
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < K; ++t) {
    if (threadIdx.x + t * T < nb) {
      pass_math<T2, N, R, FWD>(v[t], j[t], Ns, tw);
      const int d = (j[t] / Ns) * Ns * R + j[t] % Ns;
#pragma unroll
      for (int r = 0; r < R; ++r) lds[base[t] + (d + r * Ns) * L.ps] = v[t][r];
    }
  }
  __syncthreads();
}
template <typename T2, int N, bool FWD, int K, int T, int Ns, int R0, int... Rs>
__device__ __forceinline__ void chain_k(T2* lds, const Lines& L, const T2* tw) {
  pass_k<T2, N, R0, FWD, K, T>(lds, L, Ns, tw);
  if constexpr (sizeof...(Rs) > 0) chain_k<T2, N, FWD, K, T, Ns * R0, Rs...>(lds, L, tw);
}
template <typename T2, int N, bool FWD, int K, int T, int... Rs>
__device__ __forceinline__ void lines_k(T2* lds, const Lines& L, const T2* tw, Radices<Rs...>) {
  chain_k<T2, N, FWD, K, T, 1, Rs...>(lds, L, tw);
}

template <int N> constexpr size_t plane_lds() {
  return (size_t)(N + 2 * (N / 2) + N * plane_pitch(N)) * sizeof(float2);
}

// LDS of a plane work group: the N-point (y) and M-point (z) twiddles, the split factors, the plane.
template <int N> struct PlaneLds {
  static constexpr int M = N / 2, P = plane_pitch(N), Q = M / 2, T = plane_threads<N>();
  float2 *tw, *twz, *rt, *pl;
  __device__ PlaneLds(float2* smem) : tw(smem), twz(smem + N), rt(smem + N + M), pl(smem + N + 2 * M) {}
  __device__ void tables(const float2* __restrict__ twN, const float2* __restrict__ twM,
                         const float2* __restrict__ rtn) const {
    for (int i = threadIdx.x; i < N; i += T) tw[i] = twN[i];
    for (int i = threadIdx.x; i < M; i += T) {
      twz[i] = twM[i];
      rt[i] = rtn[i];
    }
  }
  // half-spectrum rows rb .. rb + n - 1 of one field (rb = x*n) <-> the plane
  __device__ void load(const float2* __restrict__ field, const Geo& g, unsigned rb) const {
    for (int e = threadIdx.x; e < N * M; e += T) pl[(e / M) * P + e % M] = at32(field, rb * M + (unsigned)e);
    for (int y = threadIdx.x; y < N; y += T) pl[y * P + M] = at32(field, (unsigned)g.nq + rb + y);
  }
  __device__ void store(float2* __restrict__ field, const Geo& g, unsigned rb) const {
    for (int e = threadIdx.x; e < N * M; e += T) at32(field, rb * M + (unsigned)e) = pl[(e / M) * P + e % M];
    for (int y = threadIdx.x; y < N; y += T) at32(field, (unsigned)g.nq + rb + y) = pl[y * P + M];
  }
  // y transform of the n/2 + 1 columns (the Nyquist column last). n <= 64: one transform of all
  // columns, up to KY butterflies per thread (32^3 step 0.0282 -> 0.0271 ms, 64^3 0.0378 ->
  // 0.0364); n = 128: the n/2 main columns, then the Nyquist column (one call of all 65 columns,
  // two butterflies per thread in the radix-8 pass: 0.1190 -> 0.1237 ms).
  static constexpr int KY = ((M + 1) * (N / PlanOf<N>::rmin) + T - 1) / T;
  template <bool FWD> __device__ void ypass() const {
    using Plan = typename PlanOf<N>::type;
    if constexpr (N <= 64) {
      using Pl = typename std::conditional<FWD, Plan, typename RevImpl<Plan, Radices<>>::type>::type;
      lines_k<float2, N, FWD, KY, T>(pl, Lines{M + 1, M + 1, 1, 0, P}, tw, Pl{});
    } else {
      fft_lines_p<float2, N, FWD, Plan>(pl, Lines{M, M, 1, 0, P}, tw);
      fft_lines_p<float2, N, FWD, Plan>(pl + M, Lines{1, 1, 0, 0, P}, tw);
    }
  }
  // C2R: split of every row into the packed M-point spectrum (task (y, k) owns the pair k, M - k;
  // k = 0 pairs X_0 with the Nyquist value X_M in slot M and, with their imaginary parts taken as
  // zero, the general formula at w = 1 gives Z_0 = (X_0 + X_M, X_0 - X_M), its second output going
  // back to slot M, which the M-point transform does not read; the self-paired M/2 of row y is
  // split by thread y), then the M-point inverse along the rows
  __device__ void c2r() const {
    for (int e = threadIdx.x; e < N * Q; e += T) {
      float2* z = pl + (e / Q) * P;
      const int k = e % Q;
      const bool k0 = k == 0;
      float2 xk = z[k], xm = z[M - k];
      xk.y = k0 ? 0.f : xk.y;
      xm.y = k0 ? 0.f : xm.y;
      float2 zk, zm;
      c2r_pair(xk, xm, rt[k], zk, zm);
      z[k] = zk;
      z[M - k] = zm;
    }
    for (int y = threadIdx.x; y < N; y += T) {
      float2* z = pl + y * P;
      const float2 c = z[Q];
      float2 zh, unused;
      c2r_pair(c, c, rt[Q], zh, unused);
      z[Q] = zh;
    }
    __syncthreads();
    fft_lines_p<float2, M, false, typename ZPlanOf<M>::type>(pl, Lines{N, N, P, 0, 1}, twz);
  }
  // R2C: M-point forward along the rows, then the split back to n/2 + 1 values per row (k = 0:
  // Z_0 with itself, X_0 = Z_0.x + Z_0.y and X_M = Z_0.x - Z_0.y into slot M; X_{M/2} = conj
  // Z_{M/2} by thread y)
  __device__ void r2c() const {
    fft_lines_p<float2, M, true, typename ZPlanOf<M>::type>(pl, Lines{N, N, P, 0, 1}, twz);
    for (int e = threadIdx.x; e < N * Q; e += T) {
      float2* z = pl + (e / Q) * P;
      const int k = e % Q;
      const bool k0 = k == 0;
      float2 xk, xm;
      r2c_pair(z[k], z[k0 ? 0 : M - k], rt[k], xk, xm);
      xk.y = k0 ? 0.f : xk.y;
      xm.y = k0 ? 0.f : xm.y;
      z[k] = xk;
      z[M - k] = xm;
    }
    for (int y = threadIdx.x; y < N; y += T) {
      float2* z = pl + y * P;
      const float2 zh = z[Q];
      z[Q] = make_float2(zh.x, -zh.y);
    }
    __syncthreads();
  }
  // the squared real rows (packed: element m of row y holds voxels 2m, 2m+1) -> y,z-transformed
  // eta^2 written to the field
  __device__ void square_out(float2* __restrict__ field, const Geo& g, unsigned rb) const {
    r2c();
    ypass<true>();
    store(field, g, rb);
  }
};

// Grid: (n planes, V variants[, members]) through geo_tile / geo_field. VT: the variant count
// (1..4). mt (may be null): batched ensemble launch, member blockIdx.z's buffers and constants.
template <int N, int VT, bool BT = false>
__global__ void __launch_bounds__(plane_threads<N>()) k_plane_mid(float2* __restrict__ kbuf, const float* __restrict__ eta,
                                                              Geo g, MartK mk, const float2* __restrict__ twN,
                                                              const float2* __restrict__ twM,
                                                              const float2* __restrict__ rtn,
                                                              const MemberPtrs* __restrict__ mt) {
  static_assert(plane_size<N>() && VT >= 1 && VT <= 4, "plane pass: n = 32, 64, 128 and 1..4 variants");
  if (BT && mt) {
    kbuf = mt[blockIdx.z].kbuf;
    eta = mt[blockIdx.z].fld;
    mk = mt[blockIdx.z].mk;
  }
  using PL = PlaneLds<N>;
  constexpr int Q = PL::Q, P = PL::P, T = PL::T;
  extern __shared__ float2 smem[];
  const PL s(smem);
  const int p = geo_field(g);
  const unsigned rb = (unsigned)geo_tile(g) * N;  // row index (x*n + y) of the plane's ky = 0
  float2* field = kbuf + p * g.kfs;
  s.load(field, g, rb);
  s.tables(twN, twM, rtn);
  __syncthreads();
  s.template ypass<false>();
  s.c2r();

  // forces on quads of voxels: packed element 2q, 2q+1 of row y holds R at voxels 4q..4q+3
  const uint32_t step = step_of(mk.step, mk.stepp);
  const float* eb = eta + (size_t)rb * N;
  for (int e = threadIdx.x; e < N * Q; e += T) {
    const int y = e / Q, q = e % Q;
    float2* z = s.pl + y * P;
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
  s.r2c();
  s.template ypass<true>();
  s.store(field, g, rb);
}

// Grid: (n planes, F fields[, members]). adv (may be null): step counter advanced once by the
// launch (per member in a batched launch, whose table also supplies kbuf and the real fields).
template <int N, bool BT = false>
__global__ void __launch_bounds__(plane_threads<N>()) k_plane_sq(float2* __restrict__ kbuf, float* __restrict__ out, Geo g,
                                                             const float2* __restrict__ twN,
                                                             const float2* __restrict__ twM,
                                                             const float2* __restrict__ rtn, uint32_t* adv,
                                                             const MemberPtrs* __restrict__ mt) {
  static_assert(plane_size<N>(), "plane pass: n = 32, 64, 128");
  if (BT && mt) {
    kbuf = mt[blockIdx.z].kbuf;
    out = mt[blockIdx.z].fld;
    adv = mt[blockIdx.z].dstep;
  }
  if (adv && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *adv += 1u;
  using PL = PlaneLds<N>;
  constexpr int Q = PL::Q, P = PL::P, T = PL::T;
  extern __shared__ float2 smem[];
  const PL s(smem);
  const int f = geo_field(g);
  const unsigned rb = (unsigned)geo_tile(g) * N;
  float2* field = kbuf + f * g.kfs;
  s.load(field, g, rb);
  s.tables(twN, twM, rtn);
  __syncthreads();
  s.template ypass<false>();
  s.c2r();
  float* dst = out + f * g.nvox + (size_t)rb * N;
  for (int e = threadIdx.x; e < N * Q; e += T) {
    const int y = e / Q, q = e % Q;
    float2* z = s.pl + y * P;
    const float2 a = z[2 * q], b = z[2 * q + 1];
    at32(reinterpret_cast<float4*>(dst), (unsigned)(y * N + 4 * q) / 4) = make_float4(a.x, a.y, b.x, b.y);
    // each task squares the very elements it has just read: no barrier between
    z[2 * q] = make_float2(a.x * a.x, a.y * a.y);
    z[2 * q + 1] = make_float2(b.x * b.x, b.y * b.y);
  }
  __syncthreads();
  s.square_out(field, g, rb);
}

// Grid: (n planes, F fields). The real fields `in` -> y,z-transformed squares in kbuf.
template <int N>
__global__ void __launch_bounds__(plane_threads<N>()) k_plane_sq0(const float* __restrict__ in, float2* __restrict__ kbuf,
                                                              Geo g, const float2* __restrict__ twN,
                                                              const float2* __restrict__ twM,
                                                              const float2* __restrict__ rtn) {
  static_assert(plane_size<N>(), "plane pass: n = 32, 64, 128");
  using PL = PlaneLds<N>;
  constexpr int Q = PL::Q, P = PL::P, T = PL::T;
  extern __shared__ float2 smem[];
  const PL s(smem);
  const int f = geo_field(g);
  const unsigned rb = (unsigned)geo_tile(g) * N;
  const float* src = in + f * g.nvox + (size_t)rb * N;
  for (int e = threadIdx.x; e < N * Q; e += T) {
    const int y = e / Q, q = e % Q;
    float2* z = s.pl + y * P;
    const float4 v = at32(reinterpret_cast<const float4*>(src), (unsigned)(y * N + 4 * q) / 4);
    z[2 * q] = make_float2(v.x * v.x, v.y * v.y);
    z[2 * q + 1] = make_float2(v.z * v.z, v.w * v.w);
  }
  s.tables(twN, twM, rtn);
  __syncthreads();
  s.square_out(kbuf + f * g.kfs, g, rb);
}

}  // namespace gsk
