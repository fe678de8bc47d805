// Strided-axis passes (y and x) of the FP32 fused step.
//
// A tile is W lines ("columns" c) of n positions along the transformed axis.
// Element (c, pos) of a field lives at field + off + c*cs + pos*ps.
//   main y tile  (x, chunk):   W consecutive kz < n/2 of one x-plane row block
//   main x tile  (ky, chunk):  W consecutive kz < n/2 of one ky
//   Nyquist tiles:             the kz = n/2 column for W consecutive x (y pass) or ky (x pass)
// Main tiles are therefore always full and every row segment they touch is a
// whole 128-byte line (W = 16 complex); the Nyquist plane, 1/(n/2+1) of the data and
// stored contiguously after the rows, is handled by n/W extra tiles.
//
// The first Stockham pass reads its operands straight from HBM and the last pass
// writes straight to HBM (fft_g2g / fft_g2r / fft_r2g), so a y pass touches LDS
// once per interior pass only, and the k-space pointwise work of the x pass runs on
// registers between the forward and the inverse transform.
#pragma once
#include "kernels_z.h"

// GSK_XUPD_PF: the martensite x-update loads its k-space state before the forward transform
#ifndef GSK_XUPD_PF
#define GSK_XUPD_PF 1
#endif
// GSK_XC_PFB: the contraction loads its B tensor before the forward transforms at the sizes
// where that was measured to pay (64, 96, 192 under a plan ending on radix 8: -7 to -12 %; 128
// and 512: slower), except where xc_cap64 holds
#ifndef GSK_XC_PFB
#define GSK_XC_PFB 1
#endif
// At n = 192 with three variants the contraction's 1024-thread groups are capped at 64 VGPRs
// (8 waves/SIMD, two groups per CU; 4 registers spill) and skip the B prefetch, which would need
// 48 more: 84.6 -> 72.5 us against one group per CU with the prefetch.
template <int N, int V> constexpr bool xc_cap64() { return N == 192 && V == 3 && gsk::PlanOf<N>::rmin == 3; }
template <int N, int V> constexpr int xc_wpe() { return xc_cap64<N, V>() ? 8 : 1; }

namespace gsk {

struct STile {
  long off, cs, ps;  // base offset, column stride, position stride (complex elements)
  int ky0, dky;      // column c has ky = ky0 + dky*c   (x tiles)
  int kz0, dkz;      //              kz = kz0 + dkz*c
};

// Tile numbering: the n/W Nyquist tiles (scattered, latency-bound) come first so they
// run while the main tiles stream, then the n*nchm main tiles.
// Main tiles read W consecutive kz of a row (one 128-byte line at W = 16); a Nyquist tile
// covers W whole rows (x or ky) of the contiguous Nyquist plane.
__device__ __forceinline__ STile ytile(int bid, const Geo& g) {
  const long M = g.M;
  const int nnyq = g.n / g.W;
  if (bid >= nnyq) {
    const int m = bid - nnyq, x = m / g.nchm, ch = m % g.nchm;
    return STile{(long)x * g.n * M + (long)ch * g.W, 1, M, 0, 0, ch * g.W, 1};
  }
  return STile{(long)g.nq + (long)bid * g.W * g.n, g.n, 1, 0, 0, g.M, 0};
}

__device__ __forceinline__ STile xtile(int t, const Geo& g) {
  const long M = g.M;
  const int nnyq = g.n / g.W;
  if (t >= nnyq) {
    const int m = t - nnyq, ky = m / g.nchm, ch = m % g.nchm;
    return STile{(long)ky * M + (long)ch * g.W, 1, (long)g.n * M, ky, 0, ch * g.W, 1};
  }
  return STile{(long)g.nq + (long)t * g.W, 1, g.n, t * g.W, 1, g.M, 0};
}

// Element (c, pos) of a tile: a uniform base plus a 32-bit offset (a field holds fewer than
// 2^29 elements), so the index math stays in 32-bit registers and the accesses use a scalar base.
struct TileIO {
  float2* p;
  unsigned cs, ps;
  __device__ __forceinline__ float2 get(int c, int pos) const { return at32(p, (unsigned)c * cs + (unsigned)pos * ps); }
  __device__ __forceinline__ void put(int c, int pos, float2 v) const { at32(p, (unsigned)c * cs + (unsigned)pos * ps) = v; }
};
__device__ __forceinline__ TileIO tio(float2* field, const STile& s) {
  return TileIO{field + s.off, (unsigned)s.cs, (unsigned)s.ps};
}

// ---------------------------------------------------------------- y pass (C2C along y, in place)
template <int N, bool FWD, bool BT = false>
__global__ void __launch_bounds__(xy_threads<N>()) k_ypass(float2* __restrict__ kbuf, Geo g, int F,
                                               const float2* __restrict__ twN, const MemberPtrs* __restrict__ mt) {
  if (BT && mt) kbuf = mt[blockIdx.z].kbuf;  // batched ensemble launch
  extern __shared__ float2 smem[];
  float2* tw = smem;
  float2* tile = smem + N;
  const int W = g.W;
  const STile s = ytile(geo_tile(g), g);
  load_tw<N>(tw, twN);
  __syncthreads();
  Lines Ls{W, W, 1, 0, W};
  // a grid with gridDim.y > 1 gives each field its own work groups
  const int fa = gridDim.y > 1 ? geo_field(g) : 0, fb = gridDim.y > 1 ? fa + 1 : F;
  for (int f = fa; f < fb; ++f) {
    const TileIO io = tio(kbuf + f * g.kfs, s);
    fft_g2g<float2, N, FWD>(
        tile, Ls, tw, [&](int c, int y) { return io.get(c, y); }, [&](int c, int y, float2 v) { io.put(c, y, v); });
  }
}

// ---------------------------------------------------------------- x pass: forward or inverse only
template <int N, bool FWD>
__global__ void __launch_bounds__(xy_threads<N>()) k_xpass(float2* __restrict__ kbuf, Geo g, int F, float scale,
                                               const float2* __restrict__ twN) {
  extern __shared__ float2 smem[];
  float2* tw = smem;
  float2* tile = smem + N;
  const int W = g.W;
  const STile s = xtile(blockIdx.x, g);
  load_tw<N>(tw, twN);
  __syncthreads();
  Lines Ls{W, W, 1, 0, W};
  for (int f = 0; f < F; ++f) {
    const TileIO io = tio(kbuf + f * g.kfs, s);
    fft_g2g<float2, N, FWD>(
        tile, Ls, tw, [&](int c, int kx) { return io.get(c, kx); },
        [&](int c, int kx, float2 v) { io.put(c, kx, make_float2(v.x * scale, v.y * scale)); });
  }
}

// ---------------------------------------------------------------- martensite contraction (V <= 4)
// S_hat -> x FFT (to registers) -> R_hat_p = sum_q B_pq S_hat_q / n^3 -> inverse x FFT.
template <int N, int V, bool BT = false>
__global__ void __launch_bounds__(xy_threads<N>()) __attribute__((amdgpu_waves_per_eu(xc_wpe<N, V>())))
k_xcontract_r(float2* __restrict__ kbuf, const float* __restrict__ Bt, Geo g,
                                                     MartK mk, const float2* __restrict__ twN,
                                                     const MemberPtrs* __restrict__ mt) {
  if (BT && mt) {  // batched ensemble launch
    kbuf = mt[blockIdx.z].kbuf;
    Bt = mt[blockIdx.z].Bt;
    mk = mt[blockIdx.z].mk;
  }
  constexpr int RL = edge_radix<N>();
  constexpr int NPQ = V * (V + 1) / 2;
  // B does not depend on the transforms: with at most 48 values per thread (last radix <= 8,
  // V <= 3) it can be fetched before them, so its loads are in flight during the forward FFTs
  constexpr bool PFB = GSK_XC_PFB && RL * NPQ <= 48 && (N == 64 || N == 96 || N == 192) && !xc_cap64<N, V>();
  extern __shared__ float2 smem[];
  float2* tw = smem;
  float2* tile = smem + N;
  const int W = g.W, T = N * W;
  const int t = blockIdx.x;
  const STile s = xtile(t, g);
  const float* Bb = Bt + (long)t * NPQ * T;  // uniform base, lane offsets below
  float bpre[PFB ? RL : 1][PFB ? NPQ : 1];
  if (PFB && threadIdx.x < W * (N / RL)) {
#pragma unroll
    for (int r = 0; r < (PFB ? RL : 1); ++r)
#pragma unroll
      for (int pq = 0; pq < (PFB ? NPQ : 1); ++pq)
        bpre[r][pq] = at32(Bb, (unsigned)(pq * T + threadIdx.x + r * (N / RL) * W));
  }
  load_tw<N>(tw, twN);
  __syncthreads();
  Lines Ls{W, W, 1, 0, W};
  float2 S[V][RL];
#pragma unroll
  for (int f = 0; f < V; ++f) {
    const TileIO io = tio(kbuf + f * g.kfs, s);
    fft_g2r<float2, N>(tile, Ls, tw, [&](int c, int kx) { return io.get(c, kx); }, S[f]);
  }
  const int b = threadIdx.x;
  if (b < W * (N / RL)) {
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      const int e = r * (N / RL) * W;
      float bv[NPQ];
#pragma unroll
      for (int pq = 0; pq < NPQ; ++pq) bv[pq] = PFB ? bpre[PFB ? r : 0][PFB ? pq : 0] : at32(Bb, (unsigned)(pq * T + b + e));
      float2 R[V];
#pragma unroll
      for (int p = 0; p < V; ++p) {
        float ax = 0.f, ay = 0.f;
#pragma unroll
        for (int q = 0; q < V; ++q) {
          const int lo = p < q ? p : q, hi = p < q ? q : p;
          const int pq = lo * V - (lo * (lo - 1)) / 2 + (hi - lo);
          ax += bv[pq] * S[q][r].x;
          ay += bv[pq] * S[q][r].y;
        }
        R[p] = make_float2(ax * mk.inv_n3, ay * mk.inv_n3);
      }
#pragma unroll
      for (int p = 0; p < V; ++p) S[p][r] = R[p];
    }
  }
#pragma unroll
  for (int p = 0; p < V; ++p) {
    const TileIO io = tio(kbuf + p * g.kfs, s);
    fft_r2g<float2, N>(tile, Ls, tw, S[p], [&](int c, int kx, float2 v) { io.put(c, kx, v); });
  }
}

// ---------------------------------------------------------------- martensite contraction (any V)
// All transformed tiles resident in LDS; used for V > 4.
template <int N>
__global__ void __launch_bounds__(xy_threads<N>()) k_xcontract(float2* __restrict__ kbuf, const float* __restrict__ Bt, Geo g,
                                                   MartK mk, const float2* __restrict__ twN) {
  extern __shared__ float2 smem[];
  float2* tw = smem;
  float2* tiles = smem + N;
  const int W = g.W, F = mk.F, T = N * W;
  const int t = blockIdx.x;
  const STile s = xtile(t, g);
  load_tw<N>(tw, twN);
  for (int f = 0; f < F; ++f) {
    const TileIO io = tio(kbuf + f * g.kfs, s);
    for (int e = threadIdx.x; e < T; e += blockDim.x) tiles[f * T + e] = io.get(e % W, e / W);
  }
  __syncthreads();
  Lines Ls{W, W, 1, 0, W};
  for (int f = 0; f < F; ++f) fft_lines<float2, N, true>(tiles + f * T, Ls, tw);
  const float* Bb = Bt + (long)t * mk.npq * T;
  for (int e = threadIdx.x; e < T; e += blockDim.x) {
    float2 S[8];
    for (int q = 0; q < F; ++q) S[q] = tiles[q * T + e];
    for (int p = 0; p < F; ++p) {
      float2 acc = make_float2(0.f, 0.f);
      for (int q = 0; q < F; ++q) {
        const int a = p < q ? p : q, c = p < q ? q : p;
        const int pq = a * F - (a * (a - 1)) / 2 + (c - a);
        const float bv = at32(Bb, (unsigned)(pq * T + e));
        acc.x += bv * S[q].x;
        acc.y += bv * S[q].y;
      }
      tiles[p * T + e] = make_float2(acc.x * mk.inv_n3, acc.y * mk.inv_n3);
    }
  }
  __syncthreads();
  for (int f = 0; f < F; ++f) fft_lines<float2, N, false>(tiles + f * T, Ls, tw);
  for (int f = 0; f < F; ++f) {
    const TileIO io = tio(kbuf + f * g.kfs, s);
    for (int e = threadIdx.x; e < T; e += blockDim.x) io.put(e % W, e / W, tiles[f * T + e]);
  }
}

// ---------------------------------------------------------------- k-space updates
// Each launch handles a group of G fields: the x transforms of the whole group are
// loaded and held in registers first (many loads in flight per work group), then the
// pointwise update reads/writes the tile-ordered state, then the group is transformed
// back. G is a compile-time constant so the register arrays never spill.

// Martensite: F_hat -> x FFT -> eta_hat = (eta_hat - dt L F_hat)/(1 + dt L kappa k^2) -> inverse x FFT,
// for the G fields f0 .. f0+G-1.
// MODE 0: step update; MODE 1: initialise the state from the forward transform of eta.
template <int N, int G, int MODE, bool BT = false>
__global__ void __launch_bounds__(xy_threads<N>()) k_xupdate_mart(float2* __restrict__ kbuf, float2* __restrict__ state, Geo g,
                                                      MartK mk, const float2* __restrict__ twN, int f0,
                                                      const MemberPtrs* __restrict__ mt) {
  if (BT && mt) {  // batched ensemble launch
    kbuf = mt[blockIdx.z].kbuf;
    state = mt[blockIdx.z].state;
    mk = mt[blockIdx.z].mk;
  }
  constexpr int RL = edge_radix<N>();
  extern __shared__ float2 smem[];
  float2* tw = smem;
  float2* tile = smem + N;
  const int W = g.W, F = mk.F, T = N * W;
  const int t = geo_tile(g);
  if (gridDim.y > 1) f0 += geo_field(g) * G;  // one group of G fields per blockIdx.y
  const STile s = xtile(t, g);
  const int b = threadIdx.x;
  const bool act = b < W * (N / RL);
  const int c = b % W, j = b / W;
  // the state this thread updates does not depend on the transform: fetch it first
  float2 sv[GSK_XUPD_PF ? G : 1][RL];
  if (GSK_XUPD_PF && MODE == 0 && act) {
#pragma unroll
    for (int gi = 0; gi < (GSK_XUPD_PF ? G : 1); ++gi) {
      const float2* st = state + ((long)t * F + f0 + gi) * T;
#pragma unroll
      for (int r = 0; r < RL; ++r) sv[gi][r] = at32(st, (unsigned)(b + r * (N / RL) * W));
    }
  }
  load_tw<N>(tw, twN);
  __syncthreads();
  Lines Ls{W, W, 1, 0, W};
  const float kyv = kcomp(s.ky0 + s.dky * c, N, mk.kfac);
  const float kzv = mk.kfac * (float)(s.kz0 + s.dkz * c);
  {
    float2 v[G][RL];
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
        const TileIO io = tio(kbuf + (f0 + gi) * g.kfs, s);
        fft_g2r<float2, N>(tile, Ls, tw, [&](int cc, int kx) { return io.get(cc, kx); }, v[gi]);
      }
    if (act) {
#pragma unroll
      for (int gi = 0; gi < G; ++gi) {
          float2* st = state + ((long)t * F + f0 + gi) * T;
#pragma unroll
          for (int r = 0; r < RL; ++r) {
            const int e = r * (N / RL) * W;
            if (MODE == 1) {
              at32(st, (unsigned)(b + e)) = v[gi][r];
            } else {
              const float kxv = kcomp(j + r * (N / RL), N, mk.kfac);
              const float k2 = kxv * kxv + kyv * kyv + kzv * kzv;
              const float den = 1.f / (1.f + mk.dtLk * k2);
              const float2 eh = GSK_XUPD_PF ? sv[GSK_XUPD_PF ? gi : 0][r] : at32(st, (unsigned)(b + e));
              const float2 nh = make_float2((eh.x - mk.dtL * v[gi][r].x) * den, (eh.y - mk.dtL * v[gi][r].y) * den);
              at32(st, (unsigned)(b + e)) = nh;
              v[gi][r] = make_float2(nh.x * mk.inv_n3, nh.y * mk.inv_n3);
            }
          }
        }
    }
    if (MODE == 0) {
#pragma unroll
      for (int gi = 0; gi < G; ++gi) {
          const TileIO io = tio(kbuf + (f0 + gi) * g.kfs, s);
          fft_r2g<float2, N>(tile, Ls, tw, v[gi], [&](int cc, int kx, float2 val) { io.put(cc, kx, val); });
        }
    }
  }
}

// Precipitate: fields (mu_c, F_1..F_V) with state (dc_hat, eta_hat_1..V).
// c: dc_hat <- (dc_hat - dt M k^2 (mu_hat + (xi B(k) - S) dc_hat)) / (1 + dt M (S k^2 + kappa_c k^4))
// eta: eta_hat <- (eta_hat - dt L F_hat) / (1 + dt L kappa_eta k^2)
// MODE 1 initialises the state; the k = 0 mode of dc is set from its FP64 sum.
// Processes the G fields f0 .. f0+G-1.
template <int N, int G, int MODE>
__global__ void __launch_bounds__(xy_threads<N>()) k_xupdate_prec(float2* __restrict__ kbuf, float2* __restrict__ state,
                                                      const float* __restrict__ Bt, Geo g, PrecK pk,
                                                      const float2* __restrict__ twN, int f0) {
  constexpr int RL = edge_radix<N>();
  extern __shared__ float2 smem[];
  float2* tw = smem;
  float2* tile = smem + N;
  const int W = g.W, F = pk.V + 1, T = N * W;
  // the fused precipitate step advances the noise counter here: the fused last z pass draws
  // the next step's noise
  if (MODE == 0 && pk.adv && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *pk.adv += 1u;
  const int t = geo_tile(g);
  if (gridDim.y > 1) f0 += geo_field(g) * G;  // one group of G fields per blockIdx.y
  const STile s = xtile(t, g);
  const int b = threadIdx.x;
  const bool act = b < W * (N / RL);
  const int c = b % W, j = b / W;
  // the state does not depend on the transform: fetch it first
  float2 sv[GSK_XUPD_PF ? G : 1][RL];
  const float* Bb = Bt + (long)t * T;
  if (GSK_XUPD_PF && MODE == 0 && act) {
#pragma unroll
    for (int gi = 0; gi < (GSK_XUPD_PF ? G : 1); ++gi) {
      const float2* st = state + ((long)t * F + f0 + gi) * T;
#pragma unroll
      for (int r = 0; r < RL; ++r) sv[gi][r] = at32(st, (unsigned)(b + r * (N / RL) * W));
    }
  }
  load_tw<N>(tw, twN);
  __syncthreads();
  Lines Ls{W, W, 1, 0, W};
  const int ky = s.ky0 + s.dky * c, kz = s.kz0 + s.dkz * c;
  const float kyv = kcomp(ky, N, pk.kfac);
  const float kzv = pk.kfac * (float)kz;
  {
    float2 v[G][RL];
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
        const TileIO io = tio(kbuf + (f0 + gi) * g.kfs, s);
        fft_g2r<float2, N>(tile, Ls, tw, [&](int cc, int kx) { return io.get(cc, kx); }, v[gi]);
      }
    if (act) {
#pragma unroll
      for (int gi = 0; gi < G; ++gi) {
          const int f = f0 + gi;
          float2* st = state + ((long)t * F + f) * T;
#pragma unroll
          for (int r = 0; r < RL; ++r) {
            const int e = r * (N / RL) * W;
            const int kx = j + r * (N / RL);
            if (MODE == 1) {
              float2 z = v[gi][r];
              if (f == 0 && kx == 0 && ky == 0 && kz == 0) z = make_float2(pk.k0, 0.f);
              at32(st, (unsigned)(b + e)) = z;
            } else {
              const float kxv = kcomp(kx, N, pk.kfac);
              const float k2 = kxv * kxv + kyv * kyv + kzv * kzv;
              const float2 sh = GSK_XUPD_PF ? sv[GSK_XUPD_PF ? gi : 0][r] : at32(st, (unsigned)(b + e));
              const float2 fh = v[gi][r];
              float2 nh;
              if (f == 0) {
                const float bv = pk.celastic ? at32(Bb, (unsigned)(b + e)) : 0.f;
                const float a = pk.xi * bv - pk.S;
                const float mx = fh.x + a * sh.x, my = fh.y + a * sh.y;
                const float den = 1.f / (1.f + pk.dtM * pk.S * k2 + pk.dtMkc * k2 * k2);
                nh = make_float2((sh.x - pk.dtM * k2 * mx) * den, (sh.y - pk.dtM * k2 * my) * den);
              } else {
                const float den = 1.f / (1.f + pk.dtLke * k2);
                nh = make_float2((sh.x - pk.dtL * fh.x) * den, (sh.y - pk.dtL * fh.y) * den);
              }
              at32(st, (unsigned)(b + e)) = nh;
              v[gi][r] = make_float2(nh.x * pk.inv_n3, nh.y * pk.inv_n3);
            }
          }
        }
    }
    if (MODE == 0) {
#pragma unroll
      for (int gi = 0; gi < G; ++gi) {
          const TileIO io = tio(kbuf + (f0 + gi) * g.kfs, s);
          fft_r2g<float2, N>(tile, Ls, tw, v[gi], [&](int cc, int kx, float2 val) { io.put(cc, kx, val); });
        }
    }
  }
}

}  // namespace gsk
