// Split-y schedule of the FP32 fused steps at n = 128, 192, 256: four launches per martensite step,
// three per precipitate step, three for the reference rules' elastic R.
//
// The y transform of an x-plane (length n) is split as n = P * m (P = 4, m = 64, 48 or 32):
//   inverse  x[P t + j] = sum_k1 w_m^(k1 t) A_j[k1],
//            A_j[k1]    = w_n^(k1 j) sum_k2 X[k1 + m k2] w_P^(k2 j)
//   forward  E_j[k1]    = sum_t x[P t + j] w_m^(-k1 t),
//            X[k1 + m k2] = sum_j w_P^(-k2 j) (w_n^(-k1 j) E_j[k1])          (w_L = exp(2 pi i / L))
// The x kernels apply the P-point stage and its twiddles in registers across lanes (the P values
// of one (x, k1, kz) sit in lanes c, c ^ Wc, c ^ 2 Wc of a wave), between their x transforms and
// the pointwise k-space work. Sub-plane j of an x-plane (the m rows y = P t + j) then needs only
// the m-point y transforms, so a work group keeps it in LDS through y inverse, the z pass and y
// forward: at n = 256, 64 rows of 129 values at pitch 137, 72 KB (two work groups per CU).
//
// Sub-plane j of x-plane x is stored in the rows x*n + j*m + k1 (k1 < m) of the half-spectrum
// layout (kz < n/2 in rows of n/2 values, the Nyquist value kz = n/2 in the Nyquist plane):
// A_j (inverse side, written by the x kernels) or E_j (forward side, written by the sub-plane
// kernels).
//
// Step: k_xcontract_split (E-form S_hat -> A-form R_hat), k_sub_mid (A-form R_hat -> E-form F_hat,
// the force on the sub-plane's rows), k_xupdate_split (E-form F_hat, k-space state -> A-form
// eta_hat), k_sub_sq (A-form eta_hat -> eta written, E-form S_hat of the next step).
// HBM bytes per voxel and variant: 8 + 2 + 8 + 4 V (eta of every variant, read through L2 by the
// V groups of a sub-plane) + 16 + 12, against 28 + 28 for the two y -> z -> y chains they replace.
#pragma once
#include "kernels_plane.h"

namespace gsk {

// Columns per sub-plane group of an x tile: 8 (a 64-byte row segment, 512-thread x work groups)
// or 4 (32 bytes, 256 threads); the x tiles that share a 128-byte line run on one XCD.
#ifndef GSK_SPLIT_WC
#define GSK_SPLIT_WC 8
#endif
// Waves per SIMD the sub-plane kernels are compiled for (0: the compiler's choice).
#ifndef GSK_SUB_WPE
#define GSK_SUB_WPE 0
#endif
#if GSK_SUB_WPE > 0
#define GSK_SUB_ATTR __attribute__((amdgpu_waves_per_eu(GSK_SUB_WPE)))
#else
#define GSK_SUB_ATTR
#endif

// Threads of a sub-plane work group: 512 (at most 128 VGPRs, so two groups, 2 x 72 KB of LDS,
// share a CU and one group's HBM phases overlap the other's LDS passes) or 1024.
#ifndef GSK_SUB_T
#define GSK_SUB_T 512
#endif

// Row pitch of the LDS sub-plane in complex values (0: the skewed row length plus one).
#ifndef GSK_SUB_PP
#define GSK_SUB_PP 0
#endif

constexpr int kSplitP = 4;  // sub-planes per x-plane
constexpr int kSubT = GSK_SUB_T;
// Threads of a sub-plane work group: kSubT at n = 256; 384 at n = 192 and 256 at n = 128 (six and
// four waves, so that each wave owns 16 main columns and 8 rows), and the work groups kept
// resident per CU (19 KB of LDS and about 80 VGPRs allow six at n = 128: 0.0971 -> 0.0960 ms
// per step against four).
template <int N> constexpr int sub_threads() { return N == 192 ? 384 : (N == 128 ? 256 : kSubT); }
template <int N> constexpr int sub_groups_per_cu() { return N == 128 ? 6 : 2; }
constexpr int kSplitWc = GSK_SPLIT_WC;  // x update (and the k-space state's tile order)
// Contraction's x transform and tile: GSK_XC_PLAN 0 (the edge plan 16 * 16 and kSplitWc columns
// per sub-plane group: 16 values of each field per thread), 1 (8 * 8 * 4 and 4 columns: 4 values
// per thread, 84 instead of 246 VGPRs at V = 3 before the LDS-parked field, but 232 instead of
// 203 us at 256^3) or 2 (the edge plan and 4 columns: 256-thread groups, four per CU, but 32-byte
// row segments: 185 -> 283 us).
#ifndef GSK_XC_PLAN
#define GSK_XC_PLAN 0
#endif
constexpr int kSplitWcC = GSK_XC_PLAN >= 1 ? 4 : kSplitWc;
template <int WC> constexpr int split_w() { return kSplitP * WC; }
static_assert(kSplitWc == 4 || kSplitWc == 8 || kSplitWc == 16, "sub-plane group width 4, 8 or 16");

template <int N> constexpr bool split_size() { return N == 256 || N == 192 || N == 128; }

// Geometry of the split schedule.
struct SGeo {
  int n, M, m;   // edge, n/2, rows per sub-plane (n / P)
  int ntile;     // x tiles of the update (kSplitWc columns per sub-plane group)
  int ntile_c;   // x tiles of the contraction (kSplitWcC)
  long long kfs, nq, nvox;
};
// x tiles with WC columns per sub-plane group: m / WC Nyquist tiles, then m * (M / WC) main tiles
template <int WC> __host__ __device__ constexpr int split_ntile(int m, int M) { return m / WC + m * (M / WC); }

__host__ __device__ constexpr int bitrev4(int j) { return ((j & 1) << 1) | (j >> 1); }

// The value v, which the compiler cannot see through: index arithmetic built on it in two places
// is computed twice instead of being kept live from the first to the second.
__device__ __forceinline__ int opaque(int v) {
  int r;
  asm volatile("v_mov_b32 %0, %1" : "=v"(r) : "v"(v));
  return r;
}

// x tile of the split schedule: W = P * WC columns, column c = j * WC + i (sub-plane j). Main
// tiles cover one k1 and the kz chunk [kz0, kz0 + WC); Nyquist tiles the kz = n/2 value of the WC
// consecutive k1 = k1_0 + i. Element (c, x) of a field: base + j * gs + i + x * ps (E / A form).
// After the lane stage, column c carries ky = k1 + m * bitrev(j) (and kz).
template <int WC> struct XSTile {
  unsigned base, gs, ps;
  int k1, kz0;
  bool nyq;
  __device__ __forceinline__ unsigned coff(int c) const { return base + (unsigned)(c / WC) * gs + (unsigned)(c % WC); }
  __device__ __forceinline__ int ky(int c, int m) const { return (nyq ? k1 + c % WC : k1) + m * bitrev4(c / WC); }
  __device__ __forceinline__ int kz(int c, int M) const { return nyq ? M : kz0 + c % WC; }
  __device__ __forceinline__ int k1c(int c) const { return nyq ? k1 + c % WC : k1; }
};

// Work group -> tile. Work groups are dealt to the 8 XCDs round robin (XCD = id % 8); the G
// tiles Gu .. Gu + G - 1 share every 128-byte line of their row segments, so the ids
// 8Gq + k + 8i (i < G: same XCD k, dispatched together) take them and each line is fetched into
// one L2.
template <int WC> __device__ __forceinline__ int xs_swizzle(int b, int ntile) {
  constexpr int G = 16 / WC, B = 8 * G;
  if (G == 1 || b >= ntile / B * B) return b;
  const int r = b % B;
  return b - r + G * (r % 8) + r / 8;
}

template <int WC> __device__ __forceinline__ XSTile<WC> xs_tile(int t, const SGeo& g) {
  const int nnyq = g.m / WC, nchm = g.M / WC;
  if (t < nnyq) return XSTile<WC>{(unsigned)(g.nq + (long long)t * WC), (unsigned)g.m, (unsigned)g.n, t * WC, 0, true};
  const int u = t - nnyq, k1 = u / nchm, ch = u % nchm;
  return XSTile<WC>{(unsigned)(k1 * g.M + ch * WC), (unsigned)(g.m * g.M), (unsigned)(g.n * g.M), k1, ch * WC, false};
}

// The P = 4 point stage across the lanes c, c ^ Wc, c ^ 2 Wc, c ^ 3 Wc (j = lane / Wc % 4).
// Forward: lane j holds w_n^(-k1 j) E_j on entry and X[k1 + m * bitrev(j)] on exit.
// Inverse: lane j holds X[k1 + m * bitrev(j)] on entry and sum_k2 X[k1 + m k2] w_4^(k2 j) on exit.
// Butterfly of the lane pairs (l, l ^ MASK) of a wave: lower lane lo + hi, upper lane lo - hi,
// as fma(sg, hi, lo) with sg = +1 on the lower and -1 on the upper lane (the product is exact, so
// this rounds as the sum or difference does). The partner value comes over the VALU: DPP row
// rotation by 8 (MASK 8, = lane ^ 8 inside a row of 16), gfx950's v_permlane16_swap /
// v_permlane32_swap (MASK 16 / 32; swapping a register with itself leaves the lower row's value
// in the first result and the upper row's in the second, in both lanes of a pair). Other masks:
// ds_bpermute through the LDS crossbar.
template <int MASK> __device__ __forceinline__ float lane_bfly(float v, float sg) {
  if constexpr (MASK == 16 || MASK == 32) {
    const unsigned u = __float_as_uint(v);
    unsigned lo, hi;
    if constexpr (MASK == 16) {
      const auto r = __builtin_amdgcn_permlane16_swap(u, u, false, false);
      lo = r[0];
      hi = r[1];
    } else {
      const auto r = __builtin_amdgcn_permlane32_swap(u, u, false, false);
      lo = r[0];
      hi = r[1];
    }
    return __builtin_fmaf(sg, __uint_as_float(hi), __uint_as_float(lo));
  } else {
    // partner p: lower lane v + p, upper lane p - v
    const float p = MASK == 8
                        ? __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x128 /* row_ror:8 */, 0xF, 0xF, true))
                        : __shfl_xor(v, MASK, 64);
    return __builtin_fmaf(sg, v, p);
  }
}
template <int MASK> __device__ __forceinline__ float2 lane_bfly2(float2 v, float sg) {
  return make_float2(lane_bfly<MASK>(v.x, sg), lane_bfly<MASK>(v.y, sg));
}
// Per-lane constants of the stage (every lane of the wave active; j = lane / Wc % 4, so j & 1 is
// lane & Wc and j & 2 lane & 2 Wc).
struct LaneStage {
  float s1, s2;  // butterfly signs of the Wc and 2 Wc steps
  bool j3;       // the lane whose value turns by -+i between the steps
  __device__ explicit LaneStage(int j) : s1((j & 1) ? -1.f : 1.f), s2((j & 2) ? -1.f : 1.f), j3(j == 3) {}
};
template <bool FWD, int WC> __device__ __forceinline__ float2 lane_dft4(float2 v, const LaneStage& ls) {
  if (FWD) {
    v = lane_bfly2<2 * WC>(v, ls.s2);
    v = make_float2(ls.j3 ? v.y : v.x, ls.j3 ? -v.x : v.y);  // rot<true> on lane 3
    v = lane_bfly2<WC>(v, ls.s1);
  } else {
    v = lane_bfly2<WC>(v, ls.s1);
    v = make_float2(ls.j3 ? -v.y : v.x, ls.j3 ? v.x : v.y);  // rot<false> on lane 3
    v = lane_bfly2<2 * WC>(v, ls.s2);
  }
  return v;
}

// x update: the edge plan (16 * 16 at n = 256, 12 * 16 at 192), kSplitWc columns per sub-plane
// group, one thread per butterfly of the plan's smallest radix. The last forward pass leaves its
// RL = edge-radix values per thread in the first W * N / RL threads (all of them at 256; 384 of
// 512 at 192, whole waves); split_active() tells a thread whether it holds them.
template <int N> constexpr int split_xthreads() { return split_w<kSplitWc>() * (N / PlanOf<N>::rmin); }
template <int N, int W, int RL, int TB> __device__ __forceinline__ bool split_active(int b) {
  static_assert((W * (N / RL)) % 64 == 0, "the holding threads fill whole waves");
  return W * (N / RL) == TB || b < W * (N / RL);
}
// contraction: the edge plan, or 8 * 8 * 4 (last radix 4: 4 values of each field per thread)
template <int N> struct XcPlan {
  using type = FwdPlan<N>;
  static constexpr int rmin = PlanOf<N>::rmin, rlast = edge_radix<N>();
};
#if GSK_XC_PLAN == 1
template <> struct XcPlan<256> {
  using type = Radices<8, 8, 4>;
  static constexpr int rmin = 4, rlast = 4;
};
#endif
template <int N> constexpr int split_cthreads() { return split_w<kSplitWcC>() * (N / XcPlan<N>::rmin); }

// B tensor in the contraction's tile order, packed p <= q: ((t * npq + pq) * n + kx) * W + c.
template <int WC>
__global__ void k_btensor_split(float* __restrict__ Bt, SGeo g, BArgs a) {
  constexpr int W = split_w<WC>();
  const int n = g.n, T = n * W;
  const int V = a.V1, npq = V * (V + 1) / 2;
  const long total = (long)split_ntile<WC>(g.m, g.M) * T;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int e = (int)(i % T);
    const long t = i / T;
    const XSTile<WC> s = xs_tile<WC>((int)t, g);
    const int c = e % W, kx = e / W;
    double out[64];
    btensor_half_spec(kx, s.ky(c, g.m), s.kz(c, g.M), n, a, out);
    for (int p = 0; p < V; ++p)
      for (int q = p; q < V; ++q) {
        const int pq = p * V - (p * (p - 1)) / 2 + (q - p);
        Bt[(t * npq + pq) * T + e] = (float)out[p * V + q];
      }
  }
}

// Contraction: E-form S_hat -> x FFT -> lane stage -> R_hat_p = sum_q B_pq S_hat_q / n^3 -> inverse
// lane stage -> inverse x FFT -> A-form R_hat (MODE 0). MODE 1: the elastic energy sum
// sum_k w(kz) sum_pq B_pq Re(S_p* S_q) of the tile into part[tile] instead (w = 1 for kz in
// {0, n/2}, 2 otherwise).
//
// MODE 0 holds V - 1 fields' spectra in registers (16 values each at n = 256) and parks the last
// field's in the LDS tile, at the positions the thread itself holds, so that at V <= 3 the kernel
// fits in 128 VGPRs: two 512-thread work groups (2 x 66 KB of LDS) per CU, 4 waves per SIMD.
// Every engine call takes the thread index through opaque(): the index arithmetic of a transform
// (16 global offsets, 16 twiddle addresses) is then recomputed per field instead of being kept live
// across all of them, and the power-of-two mask tells the compiler that every pass has all threads
// active (one basic block, so global accesses use a scalar base and one 32-bit offset register).
template <int V> constexpr int xc_waves() { return V <= 3 ? 4 : 2; }
template <int N, int V, int MODE>
__global__ void __launch_bounds__(split_cthreads<N>()) __attribute__((amdgpu_waves_per_eu(xc_waves<V>())))
k_xcontract_split(float2* __restrict__ kbuf, const float* __restrict__ Bt, SGeo g, MartK mk,
                  const float2* __restrict__ twN, double* __restrict__ part) {
  constexpr int WC = kSplitWcC, RL = XcPlan<N>::rlast, W = split_w<WC>(), T = N * W;
  constexpr int TB = split_cthreads<N>();
  static_assert((TB & (TB - 1)) == 0, "power-of-two work group");
  using Plan = typename XcPlan<N>::type;
  constexpr int NPQ = V * (V + 1) / 2;
  extern __shared__ float2 smem[];
  __shared__ double sh[16];
  float2* tw = smem;
  float2* tile = smem + N;
  const int t = xs_swizzle<WC>(blockIdx.x, g.ntile_c);
  const XSTile<WC> s = xs_tile<WC>(t, g);
  const int b = (int)threadIdx.x & (TB - 1), c = b % W, jl = c / WC;
  auto bo = [&]() { return opaque(b) & (TB - 1); };
  const LaneStage lst(jl);
  load_tw<N>(tw, twN);
  __syncthreads();
  const Lines Ls{W, W, 1, 0, W};
  // w_n^(-k1 j): the lane stage's twiddle of this column
  const float2 wt = tw[(s.k1c(c) * jl) % N];
  const float* Bb = Bt + (long)t * NPQ * T;
  if constexpr (MODE == 0) {
    constexpr int PV = V - 1;  // fields whose spectra stay in registers
    float2 S[PV > 0 ? PV : 1][RL];
#pragma unroll
    for (int f = 0; f < PV; ++f) {
      const float2* fb = kbuf + f * g.kfs;
      fft_g2r_p<float2, N, Plan>(
          tile, Ls, tw, [&](int cc, int kx) { return at32(fb, s.coff(cc) + (unsigned)kx * s.ps); }, S[f], bo());
    }
    {
      const float2* fb = kbuf + PV * g.kfs;
      fft_g2l_p<float2, N, Plan>(
          tile, Ls, tw, [&](int cc, int kx) { return at32(fb, s.coff(cc) + (unsigned)kx * s.ps); }, bo());
    }
    // position r of this thread's column segment: S[f][r] in registers, tile[b + e] for field PV
    const float2 wi = make_float2(wt.x, -wt.y);
    if (split_active<N, W, RL, TB>(b)) {
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      const int e = r * (N / RL) * W;
      float2 x[V];
#pragma unroll
      for (int f = 0; f < PV; ++f) x[f] = lane_dft4<true, WC>(cmul(S[f][r], wt), lst);
      x[PV] = lane_dft4<true, WC>(cmul(tile[b + e], wt), lst);
      float bv[NPQ];
#pragma unroll
      for (int pq = 0; pq < NPQ; ++pq) bv[pq] = at32(Bb, (unsigned)(pq * T + b + e));
#pragma unroll
      for (int p = 0; p < V; ++p) {
        float ax = 0.f, ay = 0.f;
#pragma unroll
        for (int q = 0; q < V; ++q) {
          const int lo = p < q ? p : q, hi = p < q ? q : p;
          const int pq = lo * V - (lo * (lo - 1)) / 2 + (hi - lo);
          ax += bv[pq] * x[q].x;
          ay += bv[pq] * x[q].y;
        }
        const float2 R = cmul(lane_dft4<false, WC>(make_float2(ax * mk.inv_n3, ay * mk.inv_n3), lst), wi);
        if (p < PV) S[p < PV ? p : 0][r] = R;
        else tile[b + e] = R;
      }
    }
    }
    {
      float2* fb = kbuf + PV * g.kfs;
      fft_l2g_p<float2, N, Plan>(
          tile, Ls, tw, [&](int cc, int kx, float2 v) { at32(fb, s.coff(opaque(cc) & (W - 1)) + (unsigned)kx * s.ps) = v; },
          bo());
    }
#pragma unroll
    for (int f = 0; f < PV; ++f) {
      float2* fb = kbuf + f * g.kfs;
      fft_r2g_p<float2, N, Plan>(
          tile, Ls, tw, S[f], [&](int cc, int kx, float2 v) { at32(fb, s.coff(opaque(cc) & (W - 1)) + (unsigned)kx * s.ps) = v; },
          bo());
    }
  } else {
    float2 S[V][RL];
#pragma unroll
    for (int f = 0; f < V; ++f) {
      const float2* fb = kbuf + f * g.kfs;
      fft_g2r_p<float2, N, Plan>(
          tile, Ls, tw, [&](int cc, int kx) { return at32(fb, s.coff(cc) + (unsigned)kx * s.ps); }, S[f], bo());
    }
    const int kz = s.kz(c, g.M);
    const double w = (kz == 0 || kz == g.M) ? 1.0 : 2.0;
    double acc = 0.0;
    if (split_active<N, W, RL, TB>(b)) {
#pragma unroll
    for (int f = 0; f < V; ++f)
#pragma unroll
      for (int r = 0; r < RL; ++r) S[f][r] = lane_dft4<true, WC>(cmul(S[f][r], wt), lst);
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      const int e = r * (N / RL) * W;
      double sum = 0.0;
#pragma unroll
      for (int p = 0; p < V; ++p)
#pragma unroll
        for (int q = 0; q < V; ++q) {
          const int lo = p < q ? p : q, hi = p < q ? q : p;
          const int pq = lo * V - (lo * (lo - 1)) / 2 + (hi - lo);
          sum += (double)at32(Bb, (unsigned)(pq * T + b + e)) *
                 ((double)S[p][r].x * S[q][r].x + (double)S[p][r].y * S[q][r].y);
        }
      acc += w * sum;
    }
    }
    const double rr = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = rr;
  }
}

// The x update of one tile and field (grid = tiles x fields): E-form F_hat -> x FFT -> lane stage
// -> on the tile-ordered k-space state, element by element, state <- upd(state, F_hat, k^2, e) (e:
// the element's index in a tile-ordered array: the state's, the B tensor's) -> inverse lane stage ->
// inverse x FFT -> A-form field spectrum / n^3 (MODE 0). MODE 1: the state <- init(E-form
// transform, kx, ky, kz), nothing written to the k-space buffer.
template <int N, int MODE, class Init, class Upd>
__device__ __forceinline__ void xupdate_split_tile(float2* __restrict__ kbuf, float2* __restrict__ state,
                                                   const SGeo& g, float kfac, float inv_n3,
                                                   const float2* __restrict__ twN, Init init, Upd upd) {
  constexpr int WC = kSplitWc, RL = edge_radix<N>(), W = split_w<WC>(), T = N * W;
  extern __shared__ float2 smem[];
  float2* tw = smem;
  float2* tile = smem + N;
  constexpr int TB = split_xthreads<N>();
  static_assert((TB & (TB - 1)) == 0, "power-of-two work group");
  const int t = xs_swizzle<WC>(blockIdx.x, g.ntile), f = blockIdx.y, F = gridDim.y;
  const XSTile<WC> s = xs_tile<WC>(t, g);
  const int b = (int)threadIdx.x & (TB - 1), c = b % W, j = b / W, jl = c / WC;
  const bool act = split_active<N, W, RL, TB>(b);  // holds RL values after the forward transform
  const LaneStage lst(jl);
  float2* st = state + ((long)t * F + f) * T;
  float2 sv[RL];
  if (MODE == 0 && act) {  // the state does not depend on the transform: fetched first
#pragma unroll
    for (int r = 0; r < RL; ++r) sv[r] = at32(st, (unsigned)(b + r * (N / RL) * W));
  }
  load_tw<N>(tw, twN);
  __syncthreads();
  const Lines Ls{W, W, 1, 0, W};
  float2* fb = kbuf + f * g.kfs;
  float2 v[RL];
  fft_g2r<float2, N>(tile, Ls, tw, [&](int cc, int kx) { return at32(fb, s.coff(cc) + (unsigned)kx * s.ps); }, v, b);
  const float2 wt = tw[(s.k1c(c) * jl) % N];
  const int ky = s.ky(c, g.m), kz = s.kz(c, g.M);
  if (act) {
#pragma unroll
    for (int r = 0; r < RL; ++r) v[r] = lane_dft4<true, WC>(cmul(v[r], wt), lst);
    const int bs = opaque(b) & (TB - 1);  // the state stores recompute what the loads above computed
    if (MODE == 1) {
#pragma unroll
      for (int r = 0; r < RL; ++r) at32(st, (unsigned)(bs + r * (N / RL) * W)) = init(v[r], j + r * (N / RL), ky, kz);
    } else {
      const float kyv = kcomp(ky, N, kfac);
      const float kzv = kfac * (float)kz;
#pragma unroll
      for (int r = 0; r < RL; ++r) {
        const float kxv = kcomp(j + r * (N / RL), N, kfac);
        const float k2 = kxv * kxv + kyv * kyv + kzv * kzv;
        const unsigned e = (unsigned)(bs + r * (N / RL) * W);
        const float2 nh = upd(sv[r], v[r], k2, e);
        at32(st, e) = nh;
        v[r] = make_float2(nh.x * inv_n3, nh.y * inv_n3);
      }
      const float2 wi = make_float2(wt.x, -wt.y);
#pragma unroll
      for (int r = 0; r < RL; ++r) v[r] = cmul(lane_dft4<false, WC>(v[r], lst), wi);
    }
  }
  if (MODE == 1) return;
  fft_r2g<float2, N>(
      tile, Ls, tw, v, [&](int cc, int kx, float2 val) { at32(fb, s.coff(opaque(cc) & (W - 1)) + (unsigned)kx * s.ps) = val; },
      opaque(b) & (TB - 1));
}

// Martensite: eta_hat <- (eta_hat - dt L F_hat) / (1 + dt L kappa k^2). MODE 1: the state from the
// E-form transform of eta.
template <int N, int MODE>
__global__ void __launch_bounds__(split_xthreads<N>()) k_xupdate_split(float2* __restrict__ kbuf,
                                                                       float2* __restrict__ state, SGeo g, MartK mk,
                                                                       const float2* __restrict__ twN) {
  xupdate_split_tile<N, MODE>(
      kbuf, state, g, mk.kfac, mk.inv_n3, twN, [](float2 z, int, int, int) { return z; },
      [&](float2 sh, float2 fh, float k2, unsigned) {
        const float den = 1.f / (1.f + mk.dtLk * k2);
        return make_float2((sh.x - mk.dtL * fh.x) * den, (sh.y - mk.dtL * fh.y) * den);
      });
}

// Precipitate, field f of (dc, eta_1..V):
//   f = 0: dc_hat <- (dc_hat - dt M k^2 (mu_hat + (xi B(k) - S) dc_hat)) / (1 + dt M (S k^2 + kappa_c k^4))
//   f > 0: eta_hat <- (eta_hat - dt L F_hat) / (1 + dt L kappa_eta k^2)
// B(k) (celastic): the concentration's scalar tensor in this kernel's tile order. MODE 1: the
// state from the E-form transform of the fields, the k = 0 mode of dc set from its FP64 sum
// (pk.k0). pk.adv (MODE 0, may be null): the step counter the launch advances once.
template <int N, int MODE>
__global__ void __launch_bounds__(split_xthreads<N>()) k_xupdate_split_prec(float2* __restrict__ kbuf,
                                                                            float2* __restrict__ state,
                                                                            const float* __restrict__ Bt, SGeo g,
                                                                            PrecK pk, const float2* __restrict__ twN) {
  if (MODE == 0 && pk.adv && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *pk.adv += 1u;
  const int f = blockIdx.y;
  const float* Bb = Bt + (long)xs_swizzle<kSplitWc>(blockIdx.x, g.ntile) * (N * split_w<kSplitWc>());
  xupdate_split_tile<N, MODE>(
      kbuf, state, g, pk.kfac, pk.inv_n3, twN,
      [&](float2 z, int kx, int ky, int kz) {
        return (f == 0 && kx == 0 && ky == 0 && kz == 0) ? make_float2(pk.k0, 0.f) : z;
      },
      [&](float2 sh, float2 fh, float k2, unsigned e) {
        if (f == 0) {
          const float bv = pk.celastic ? at32(Bb, e) : 0.f;
          const float a = pk.xi * bv - pk.S;
          const float mx = fh.x + a * sh.x, my = fh.y + a * sh.y;
          const float den = 1.f / (1.f + pk.dtM * pk.S * k2 + pk.dtMkc * k2 * k2);
          return make_float2((sh.x - pk.dtM * k2 * mx) * den, (sh.y - pk.dtM * k2 * my) * den);
        }
        const float den = 1.f / (1.f + pk.dtLke * k2);
        return make_float2((sh.x - pk.dtL * fh.x) * den, (sh.y - pk.dtL * fh.y) * den);
      });
}

// Precipitate elastic energy sum_k w(kz) B(k) |dc_hat(k)|^2 from the split-order state (field 0;
// B in the update's tile order).
template <int = 0>
__global__ void r_state_el_prec_split(const float2* __restrict__ state, const float* __restrict__ Bt, SGeo g, int F,
                                      double* __restrict__ part) {
  __shared__ double sh[16];
  constexpr int W = split_w<kSplitWc>();
  const int T = g.n * W;
  const long total = (long)g.ntile * T;
  double acc = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int e = (int)(i % T);
    const long t = i / T;
    const XSTile<kSplitWc> s = xs_tile<kSplitWc>((int)t, g);
    const int kz = s.kz(e % W, g.M);
    const double w = (kz == 0 || kz == g.M) ? 1.0 : 2.0;
    const float2 v = state[(t * F + 0) * T + e];
    acc += w * (double)Bt[t * T + e] * ((double)v.x * v.x + (double)v.y * v.y);
  }
  const double r = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// Gradient energy from the split-order state: sum_f sum_k w(kz) k^2 |eta_hat_f|^2 per field.
template <int = 0>
__global__ void r_state_grad_split(const float2* __restrict__ state, SGeo g, int F, float kfac,
                                   double* __restrict__ part) {
  __shared__ double sh[16];
  constexpr int W = split_w<kSplitWc>();
  const int n = g.n, T = n * W;
  const long total = (long)g.ntile * F * T;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int e = (int)(i % T);
    const long tf = i / T;
    const int f = (int)(tf % F);
    const XSTile<kSplitWc> s = xs_tile<kSplitWc>((int)(tf / F), g);
    const int c = e % W, kx = e / W;
    const int ky = s.ky(c, g.m), kz = s.kz(c, g.M);
    const double kxv = (double)kfac * ((kx < n / 2) ? kx : kx - n), kyv = (double)kfac * ((ky < n / 2) ? ky : ky - n),
                 kzv = (double)kfac * kz;
    const double w = (kz == 0 || kz == g.M) ? 1.0 : 2.0;
    const float2 v = state[i];
    acc[f] += w * (kxv * kxv + kyv * kyv + kzv * kzv) * ((double)v.x * v.x + (double)v.y * v.y);
  }
  for (int f = 0; f < F; ++f) {
    const double r = block_sum(acc[f], sh);
    if (threadIdx.x == 0) part[blockIdx.x * F + f] = r;
  }
}

// ---------------------------------------------------------------- sub-plane kernels
// LDS of a sub-plane work group: the m-point (y) and M-point (z) twiddles, the split factors and
// the sub-plane: m rows of n/2 + 1 values. Value p of a row sits at sk(p) = p + p / 16 (pitch
// sk(n/2) + 1): the first pass of a row transform writes with stride R (8 or 16), which without
// the skew puts 16 lanes on two of the 16 LDS bank pairs; contiguous runs inside a 16-block keep
// one skew and stay conflict-free.
//
// Each wave owns CW columns and RW rows of the sub-plane. A y transform (along a column) or a row
// operation (C2R / R2C split, z transform, pointwise work) only touches the calling wave's
// columns or rows, so it needs no work-group barrier: a step has four (after the load, after y
// inverse, after the row phase, after y forward) instead of two per LDS pass.
template <int N> struct SubLds {
  __host__ __device__ static constexpr int sk(int p) { return p + (p >> 4); }
  static constexpr int M = N / 2, Y = N / kSplitP, PP = GSK_SUB_PP > 0 ? GSK_SUB_PP : sk(N / 2) + 1, Q = M / 2,
                       T = sub_threads<N>();
  static_assert(PP > sk(N / 2), "pitch holds a skewed row");
  static constexpr int NW = T / 64, RW = Y / NW, CW = M / NW;
  static_assert(Y % NW == 0 && M % NW == 0, "rows and columns in equal shares per wave");
  static_assert(CW <= 16 && 16 % CW == 0, "a wave's columns inside one 16-block (one skew)");
  float2 *twy, *twz, *rt, *pl;
  int w, lane;
  static constexpr int TZ = sk(M - 1) + 1;  // the skewed z twiddle table
  __device__ SubLds(float2* smem)
      : twy(smem), twz(smem + Y), rt(smem + Y + TZ), pl(smem + Y + TZ + M), w((int)threadIdx.x / 64),
        lane((int)threadIdx.x % 64) {}
  __device__ void tables(const float2* __restrict__ twN, const float2* __restrict__ twM,
                         const float2* __restrict__ rtn) const {
    for (int i = threadIdx.x; i < Y; i += T) twy[i] = twN[i * kSplitP];  // exp(-2 pi i k / m)
    for (int i = threadIdx.x; i < M; i += T) {
      twz[sk(i)] = twM[i];
      rt[i] = rtn[i];
    }
  }
  // rows rb .. rb + m - 1 of one field's half spectrum -> the sub-plane, in two halves: fetch()
  // into registers (the loads stay in flight while the work group transforms the previous
  // sub-plane), put() into LDS; store() writes the sub-plane back (whole work group)
  static constexpr int LE = Y * M / T;
  static_assert(Y * M % T == 0 && Y <= T, "whole rows per thread share");
  struct Pre {
    float2 v[LE], nq;
  };
  __device__ void fetch(const float2* __restrict__ field, const SGeo& g, unsigned rb, Pre& p) const {
#pragma unroll
    for (int u = 0; u < LE; ++u) p.v[u] = at32(field, rb * M + (unsigned)(threadIdx.x + u * T));
    if ((int)threadIdx.x < Y) p.nq = at32(field, (unsigned)g.nq + rb + threadIdx.x);
  }
  __device__ void put(const Pre& p) const {
#pragma unroll
    for (int u = 0; u < LE; ++u) {
      const int e = threadIdx.x + u * T;
      pl[(e / M) * PP + sk(e % M)] = p.v[u];
    }
    if ((int)threadIdx.x < Y) pl[threadIdx.x * PP + sk(M)] = p.nq;
  }
  __device__ void store(float2* __restrict__ field, const SGeo& g, unsigned rb) const {
    for (int e = threadIdx.x; e < Y * M; e += T) at32(field, rb * M + (unsigned)e) = pl[(e / M) * PP + sk(e % M)];
    for (int y = threadIdx.x; y < Y; y += T) at32(field, (unsigned)g.nq + rb + y) = pl[y * PP + sk(M)];
  }
  // m-point transforms of the wave's CW main columns (wave 0: also of the Nyquist column)
  template <bool FWD> __device__ void ypass() const {
    using Plan = typename PlanOf<Y>::type;
    const int ln = opaque(lane) & 63;  // the mask keeps the lane index known non-negative
    fft_wlines<float2, Y, FWD, Plan, CW, false>(pl + sk(w * CW), Lines{CW, CW, 1, 0, PP}, twy, ln);
    if (w == 0) fft_wlines<float2, Y, FWD, Plan, 1, false>(pl + sk(M), Lines{1, 1, 0, 0, PP}, twy, ln);
  }
  // the wave's first row
  __device__ float2* rows() const { return pl + w * RW * PP; }
  // C2R split of the wave's rows into the packed M-point spectrum, then the M-point inverse.
  // Lane k (one per row and value, Q = 64 = the lane count) splits the pair (k, M - k); lane 0's
  // partner is the Nyquist value in slot M, and its split, with the imaginary parts of X_0 and X_M
  // taken as zero, is (X_0 + X_M, X_0 - X_M): the general formula at w = 1, so no lane branches.
  // Lane 0 also writes its Z_{M-k} into slot M, which the M-point transform does not read. The
  // self-paired value k = Q of the wave's RW rows is split afterwards, one row per lane.
  __device__ void c2r() const {
    static_assert(Q <= 64 && RW <= 64, "at most one pair per lane and row");
    float2* R0 = rows();
    const int ln = opaque(lane) & 63;  // the mask keeps the lane index known non-negative
    const float2 w = rt[ln];
    const bool k0 = ln == 0;
    if (Q == 64 || ln < Q)  // lanes Q.. (n = 192) have no pair
    for (int i = 0; i < RW; ++i) {
      float2* z = R0 + i * PP;
      float2 xk = z[sk(ln)], xm = z[sk(M - ln)];
      xk.y = k0 ? 0.f : xk.y;
      xm.y = k0 ? 0.f : xm.y;
      float2 zk, zm;
      c2r_pair(xk, xm, w, zk, zm);
      z[sk(ln)] = zk;
      z[sk(M - ln)] = zm;
    }
    if (ln < RW) {
      float2* z = R0 + ln * PP;
      const float2 c = z[sk(Q)];
      float2 zh, unused;
      c2r_pair(c, c, rt[Q], zh, unused);
      z[sk(Q)] = zh;
    }
    wave_lds_sync();
    fft_wlines<float2, M, false, typename PlanOf<M>::type, RW, true, true>(R0, Lines{RW, RW, PP, 0, 1}, twz, ln);
  }
  // M-point forward along the wave's rows, then the split back to n/2 + 1 values per row: lane k
  // turns (Z_k, Z_{M-k}) into (X_k, X_{M-k}); lane 0's partner is Z_0 itself, which the general
  // formula at w = 1 turns into X_0 = Z_0.x + Z_0.y and, in slot M, X_M = Z_0.x - Z_0.y (imaginary
  // parts zero). X_Q = conj Z_Q afterwards, one row per lane.
  __device__ void r2c() const {
    float2* R0 = rows();
    const int ln = opaque(lane) & 63;  // the mask keeps the lane index known non-negative
    fft_wlines<float2, M, true, typename PlanOf<M>::type, RW, true, true>(R0, Lines{RW, RW, PP, 0, 1}, twz, ln);
    const float2 w = rt[ln];
    const bool k0 = ln == 0;
    int pm;  // the partner M - k, and 0 for k = 0
    if constexpr ((M & (M - 1)) == 0) pm = (M - ln) & (M - 1);
    else pm = k0 ? 0 : M - ln;
    if (Q == 64 || ln < Q)  // lanes Q.. (n = 192) have no pair
    for (int i = 0; i < RW; ++i) {
      float2* z = R0 + i * PP;
      const float2 zk = z[sk(ln)], zm = z[sk(pm)];
      float2 xk, xm;
      r2c_pair(zk, zm, w, xk, xm);
      xk.y = k0 ? 0.f : xk.y;
      xm.y = k0 ? 0.f : xm.y;
      z[sk(ln)] = xk;
      z[sk(M - ln)] = xm;
    }
    if (ln < RW) {
      float2* z = R0 + ln * PP;
      const float2 zh = z[sk(Q)];
      z[sk(Q)] = make_float2(zh.x, -zh.y);
    }
  }
};

template <int N> constexpr size_t sub_lds() {
  using S = SubLds<N>;
  return (size_t)(S::Y + S::TZ + S::M + S::Y * S::PP) * sizeof(float2);
}

// Middle of a step, per (sub-plane, variant p): A-form R_hat_p -> y inverse -> C2R -> G on the
// sub-plane's rows y = P t + j: F_p = A eta_p - B eta_p^2 + C eta_p sum_q eta_q^2 + 2 w eta_p R_p
// + a xi_p (eta of every variant as 16-byte quads, xi from the Philox blocks of the voxel quads)
// -> R2C -> y forward -> E-form F_hat_p.
// Resident work groups (two per CU, grid G a multiple of 8V): group b takes the items b, b + G,
// ... and loads each next item's spectrum while it transforms the current one. Item 8Vq + 8p + k
// is variant p of sub-plane 8q + k, so the V items of a sub-plane, which read the same eta, run at
// the same time on one XCD (item % 8 = b % 8 = k) and share its L2.
template <int N, int VT>
__global__ void __launch_bounds__(sub_threads<N>()) GSK_SUB_ATTR k_sub_mid(float2* __restrict__ kbuf, const float* __restrict__ eta, SGeo g,
                                                  MartK mk, const float2* __restrict__ twN,
                                                  const float2* __restrict__ twM, const float2* __restrict__ rtn) {
  using SL = SubLds<N>;
  constexpr int Q = SL::Q, PP = SL::PP, Y = SL::Y, RW = SL::RW, NI = N * kSplitP * VT;
  extern __shared__ float2 smem[];
  const SL s(smem);
  auto item = [&](int it, int& x, int& j, int& p) {
    const int grp = 8 * VT, rem = it % grp, sp = it / grp * 8 + rem % 8;
    p = rem / 8;
    x = sp / kSplitP;
    j = sp % kSplitP;
  };
  const uint32_t step = step_of(mk.step, mk.stepp);
  typename SL::Pre pre;
  {
    int x, j, p;
    item(blockIdx.x, x, j, p);
    s.fetch(kbuf + p * g.kfs, g, (unsigned)(x * N + j * Y), pre);
  }
  s.tables(twN, twM, rtn);
  for (int it = blockIdx.x; it < NI; it += gridDim.x) {
    int x, j, p;
    item(it, x, j, p);
    const unsigned rb = (unsigned)(x * N + j * Y);
    float2* field = kbuf + p * g.kfs;
    s.put(pre);
    __syncthreads();
    if (it + (int)gridDim.x < NI) {
      int x2, j2, p2;
      item(it + gridDim.x, x2, j2, p2);
      s.fetch(kbuf + p2 * g.kfs, g, (unsigned)(x2 * N + j2 * Y), pre);
    }
    s.template ypass<false>();
    __syncthreads();
    s.c2r();
    const float* eb = eta + (size_t)x * N * N;  // the x-plane's voxels
    // One voxel quad per lane and row (Q = 64). The quads of all variants of a row are requested
    // before any is used, and the variant p's quad is picked by selects, not a branch: without
    // that, each load waited for the one before it (three round trips per row).
    static_assert(Q <= 64, "at most one voxel quad per lane and row");
    const int q = s.lane;
    if (Q == 64 || q < Q)  // lanes Q.. (n = 192) hold no quad
    for (int i = 0; i < RW; ++i) {
      const int t = s.w * RW + i;
      float2* z = s.rows() + i * PP;
      const unsigned vo = (unsigned)((kSplitP * t + j) * N + 4 * q);  // voxel offset inside the x-plane
      float4 ep = make_float4(0.f, 0.f, 0.f, 0.f), s2 = make_float4(0.f, 0.f, 0.f, 0.f);
      {
        float4 tq[VT];
#pragma unroll
        for (int f = 0; f < VT; ++f) tq[f] = at32(reinterpret_cast<const float4*>(eb + (size_t)f * g.nvox), vo / 4);
#pragma unroll
        for (int f = 0; f < VT; ++f) {
          s2.x += tq[f].x * tq[f].x; s2.y += tq[f].y * tq[f].y; s2.z += tq[f].z * tq[f].z; s2.w += tq[f].w * tq[f].w;
          const bool sel = f == p;
          ep = make_float4(sel ? tq[f].x : ep.x, sel ? tq[f].y : ep.y, sel ? tq[f].z : ep.z, sel ? tq[f].w : ep.w);
        }
      }
      {
        float xi[4] = {0.f, 0.f, 0.f, 0.f};
        if (mk.noise != 0.f) normal4(mk.seed, (unsigned)((x * N * N + vo) >> 2), (uint32_t)p, step, xi);
        const float2 ra = z[SL::sk(2 * q)], rbv = z[SL::sk(2 * q + 1)];
        const float e4[4] = {ep.x, ep.y, ep.z, ep.w}, s4[4] = {s2.x, s2.y, s2.z, s2.w}, r4[4] = {ra.x, ra.y, rbv.x, rbv.y};
        float o[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const float et = e4[u];
          o[u] = mk.A * et - mk.B * et * et + mk.C * et * s4[u] + mk.two_wel * et * r4[u] + mk.noise * xi[u];
        }
        z[SL::sk(2 * q)] = make_float2(o[0], o[1]);
        z[SL::sk(2 * q + 1)] = make_float2(o[2], o[3]);
      }
    }
    wave_lds_sync();
    s.r2c();
    __syncthreads();
    s.template ypass<true>();
    __syncthreads();
    s.store(field, g, rb);
    __syncthreads();  // the next put overwrites what store reads
  }
}

// End of a step, per (sub-plane, field): A-form eta_hat -> y inverse -> C2R -> eta written to the
// real field and squared -> R2C -> y forward -> E-form S_hat of the next step. adv (may be null):
// the step counter the launch advances once. Resident work groups as k_sub_mid; item
// f * n * P + (sub-plane) is field f of that sub-plane. SQ = false: the inverse only (y inverse,
// C2R, the real field written), for fields that end in real space (the reference rule's R).
template <int N, bool SQ = true>
__global__ void __launch_bounds__(sub_threads<N>()) GSK_SUB_ATTR k_sub_sq(float2* __restrict__ kbuf, float* __restrict__ out, SGeo g,
                                                 const float2* __restrict__ twN, const float2* __restrict__ twM,
                                                 const float2* __restrict__ rtn, uint32_t* adv, int nitems) {
  if (adv && blockIdx.x == 0 && threadIdx.x == 0) *adv += 1u;
  using SL = SubLds<N>;
  constexpr int Q = SL::Q, PP = SL::PP, Y = SL::Y, RW = SL::RW, NSP = N * kSplitP;
  extern __shared__ float2 smem[];
  const SL s(smem);
  typename SL::Pre pre;
  {
    const int it = blockIdx.x, sp = it % NSP;
    s.fetch(kbuf + (it / NSP) * g.kfs, g, (unsigned)((sp / kSplitP) * N + (sp % kSplitP) * Y), pre);
  }
  s.tables(twN, twM, rtn);
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const int f = it / NSP, sp = it % NSP, x = sp / kSplitP, j = sp % kSplitP;
    const unsigned rb = (unsigned)(x * N + j * Y);
    float2* field = kbuf + f * g.kfs;
    s.put(pre);
    __syncthreads();
    const int nx = it + (int)gridDim.x;
    if (nx < nitems) {
      const int sp2 = nx % NSP;
      s.fetch(kbuf + (nx / NSP) * g.kfs, g, (unsigned)((sp2 / kSplitP) * N + (sp2 % kSplitP) * Y), pre);
    }
    s.template ypass<false>();
    __syncthreads();
    s.c2r();
    float* dst = out + f * g.nvox + (size_t)x * N * N;
    for (int i = 0; i < RW; ++i) {
      const int t = s.w * RW + i;
      float2* z = s.rows() + i * PP;
      for (int q = s.lane; q < Q; q += 64) {
        const float2 a = z[SL::sk(2 * q)], b = z[SL::sk(2 * q + 1)];
        at32(reinterpret_cast<float4*>(dst), (unsigned)((kSplitP * t + j) * N + 4 * q) / 4) = make_float4(a.x, a.y, b.x, b.y);
        if (SQ) {  // each lane squares the very elements it has just read
          z[SL::sk(2 * q)] = make_float2(a.x * a.x, a.y * a.y);
          z[SL::sk(2 * q + 1)] = make_float2(b.x * b.x, b.y * b.y);
        }
      }
    }
    if (SQ) {
      wave_lds_sync();
      s.r2c();
      __syncthreads();
      s.template ypass<true>();
      __syncthreads();
      s.store(field, g, rb);
    }
    __syncthreads();  // the next put overwrites what store (or the row phase) reads
  }
}

// Real fields `in` (OP 1: squared) -> R2C along z -> y forward -> E-form spectra in kbuf, per
// (sub-plane, field): the carried S_hat of a fresh start, bit-identical to what k_sub_sq leaves
// (same FP32 values, same code), and (OP 0) the transform of eta that initialises the state.
template <int N, int OP>
__global__ void __launch_bounds__(sub_threads<N>()) GSK_SUB_ATTR k_sub_fwd(const float* __restrict__ in, float2* __restrict__ kbuf, SGeo g,
                                                  const float2* __restrict__ twN, const float2* __restrict__ twM,
                                                  const float2* __restrict__ rtn) {
  using SL = SubLds<N>;
  constexpr int Q = SL::Q, PP = SL::PP, Y = SL::Y, RW = SL::RW;
  extern __shared__ float2 smem[];
  const SL s(smem);
  const int x = blockIdx.x / kSplitP, j = blockIdx.x % kSplitP, f = blockIdx.y;
  const unsigned rb = (unsigned)(x * N + j * Y);
  const float* src = in + f * g.nvox + (size_t)x * N * N;
  s.tables(twN, twM, rtn);
  for (int i = 0; i < RW; ++i) {  // the wave's rows
    const int t = s.w * RW + i;
    float2* z = s.rows() + i * PP;
    for (int q = s.lane; q < Q; q += 64) {
      const float4 v = at32(reinterpret_cast<const float4*>(src), (unsigned)((kSplitP * t + j) * N + 4 * q) / 4);
      if (OP == 1) {
        z[SL::sk(2 * q)] = make_float2(v.x * v.x, v.y * v.y);
        z[SL::sk(2 * q + 1)] = make_float2(v.z * v.z, v.w * v.w);
      } else {
        z[SL::sk(2 * q)] = make_float2(v.x, v.y);
        z[SL::sk(2 * q + 1)] = make_float2(v.z, v.w);
      }
    }
  }
  __syncthreads();  // the tables
  s.r2c();
  __syncthreads();
  s.template ypass<true>();
  __syncthreads();
  s.store(kbuf + f * g.kfs, g, rb);
}

// Local precipitate forces, per (sub-plane, field f): the FT = V + 1 real fields of the
// sub-plane's rows (concentration fluctuation, eta_1..V) as 16-byte voxel quads -> mu_c (f = 0)
// or F_f (prec_local4, Philox normals of the voxel quads) -> R2C along z -> y forward -> E-form
// F_hat_f in kbuf. R (may be null): the reference rule's real elastic R of the order parameters,
// added to F_f as el R_f eta_f; cnoise: the reference rule's noise on mu_c. Item 8 FT q + 8 f + k
// is field f of sub-plane 8q + k, so the FT items of a sub-plane, which read the same voxels, run
// at about the same time on one XCD (work-group index mod 8) and share its L2.
template <int N, int FT>
__global__ void __launch_bounds__(sub_threads<N>()) GSK_SUB_ATTR k_sub_local(const float* __restrict__ fld, float2* __restrict__ kbuf,
                                                    SGeo g, PrecK pk, const float* __restrict__ R, float el,
                                                    float cnoise, const float2* __restrict__ twN,
                                                    const float2* __restrict__ twM, const float2* __restrict__ rtn) {
  using SL = SubLds<N>;
  constexpr int Q = SL::Q, PP = SL::PP, Y = SL::Y, RW = SL::RW;
  static_assert(Q <= 64, "at most one voxel quad per lane and row");
  extern __shared__ float2 smem[];
  const SL s(smem);
  const int it = blockIdx.x, rem = it % (8 * FT), sp = it / (8 * FT) * 8 + rem % 8, f = rem / 8;
  const int x = sp / kSplitP, j = sp % kSplitP;
  const unsigned rb = (unsigned)(x * N + j * Y);
  const float* xb = fld + (size_t)x * N * N;  // the x-plane's voxels
  const unsigned nv = (unsigned)g.nvox;
  const uint32_t step = step_of(pk.step, pk.stepp);
  s.tables(twN, twM, rtn);
  const int q = s.lane;
  if (Q == 64 || q < Q)  // lanes Q.. (n = 192) hold no quad
  for (int i = 0; i < RW; ++i) {  // the wave's rows
    const int t = s.w * RW + i;
    float2* z = s.rows() + i * PP;
    const unsigned vo = (unsigned)((kSplitP * t + j) * N + 4 * q);  // voxel offset inside the x-plane
    float e[FT][4];
#pragma unroll
    for (int u = 0; u < FT; ++u) {  // 32-bit offsets: FT fields hold fewer than 2^31 values
      const float4 v = at32(reinterpret_cast<const float4*>(xb), ((unsigned)u * nv + vo) / 4);
      e[u][0] = v.x; e[u][1] = v.y; e[u][2] = v.z; e[u][3] = v.w;
    }
    const uint32_t gq = (uint32_t)(((unsigned)x * N * N + vo) >> 2);
    float xi[4] = {0.f, 0.f, 0.f, 0.f};
    float rk[4] = {0.f, 0.f, 0.f, 0.f};
    if (f == 0) {
      if (cnoise != 0.f) normal4(pk.seed, gq, 0x8000u, step, xi);
    } else {
      if (pk.noise != 0.f) normal4(pk.seed, gq, (uint32_t)(f - 1), step, xi);
      if (R) {
        const float4 rq = at32(reinterpret_cast<const float4*>(R + (size_t)x * N * N), ((unsigned)(f - 1) * nv + vo) / 4);
        rk[0] = rq.x; rk[1] = rq.y; rk[2] = rq.z; rk[3] = rq.w;
      }
    }
    float o[4];
    prec_local4<FT>(pk, f, e, xi, cnoise, R != nullptr, rk, el, o);
    z[SL::sk(2 * q)] = make_float2(o[0], o[1]);
    z[SL::sk(2 * q + 1)] = make_float2(o[2], o[3]);
  }
  __syncthreads();  // the tables
  s.r2c();
  __syncthreads();
  s.template ypass<true>();
  __syncthreads();
  s.store(kbuf + f * g.kfs, g, rb);
}

}  // namespace gsk
