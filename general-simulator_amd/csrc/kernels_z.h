// Contiguous-axis (z) passes of the FP32 fused step: real rows <-> half-spectrum rows.
//
// A z tile is L = zrows(M) consecutive rows (x*n + y) of one field (zmid_rows(M) in the
// middle martensite pass). The complex
// M-point FFT (M = n/2) of the packed real row runs in LDS (row pitch zpitch(M),
// compile-time tile shape and pass strides). The R2C / C2R split between the
// packed M-point spectrum and the n/2+1 half-spectrum values is fused with the
// global store / load: task (l, k), k < M/2, of a tile owns the spectrum pair
// (k, M-k) of row l (task k = 0 owns 0, M/2 and M), and task (l, q) of the real
// side owns the four voxels 4q..4q+3 of row l (one 16-byte access). Both task
// sets have L*M/2 members, so one thread handles the same (l, k = q) slots on
// both sides.
//
// Every thread issues all of its global loads for a tile before it touches LDS,
// so a work group keeps L*M*8 bytes in flight while it waits; one field per LDS
// tile keeps the footprint at about (L*zpitch(M) + 2M)*8 bytes so several work
// groups share a CU. The middle martensite pass walks the variants one after the
// other over the same LDS tile and prefetches the next variant's spectrum while
// the current one is transformed.
#pragma once
#include <type_traits>
#include "kernels_common.h"

namespace gsk {

// Compile-time shape of the z kernels. GSK_ZROWS_{128,96,64,32}: rows per tile at M = 128, 96, 64, 32;
// GSK_ZPLAN_128: radix plan of the 128-point transform (0: 16x8, 1: 8x4x4);
// GSK_ZMID_WPE: waves-per-SIMD floor of the middle pass (0: the per-shape floor of zmid_wpe).
#ifndef GSK_ZROWS_128
#define GSK_ZROWS_128 8
#endif
#ifndef GSK_ZROWS_64
#define GSK_ZROWS_64 16  // 8 rows: no change at 128^3
#endif
#ifndef GSK_ZROWS_32
#define GSK_ZROWS_32 16  // 64^3 step 0.0535 -> 0.0518 ms against 32 rows
#endif
#ifndef GSK_ZROWS_96
#define GSK_ZROWS_96 8  // 192^3 step 0.418 -> 0.406 ms against 16 rows
#endif
#ifndef GSK_ZPLAN_128
#define GSK_ZPLAN_128 1
#endif
// Rows per tile of the middle martensite pass at M = 256, 128, 64 (the other sizes use zrows):
// 512^3 middle pass 1.26 ms at 8 rows, 1.10 at 4, 1.09 at 2; 256^3 141, 137, 135 us;
// 128^3 step 0.1213 -> 0.1202 ms at 8 rows against 16; 4 rows at 192^3 and 8 at 64^3: no gain.
#ifndef GSK_ZMID_ROWS_256
#define GSK_ZMID_ROWS_256 2
#endif
#ifndef GSK_ZMID_ROWS_128
#define GSK_ZMID_ROWS_128 2
#endif
#ifndef GSK_ZMID_ROWS_64
#define GSK_ZMID_ROWS_64 8
#endif
#ifndef GSK_ZMID_WPE
#define GSK_ZMID_WPE 0
#endif

template <int M> struct ZPlanOf {
  using type = typename PlanOf<M>::type;
  static constexpr int rmin = PlanOf<M>::rmin;
};
#if GSK_ZPLAN_128 == 1
template <> struct ZPlanOf<128> {
  using type = Radices<8, 4, 4>;
  static constexpr int rmin = 4;
};
#endif
// a radix-4 pass halves the tasks per thread (fewer live registers in the middle pass)
template <> struct ZPlanOf<256> {
  using type = Radices<4, 8, 8>;
  static constexpr int rmin = 4;
};

// Plan of the middle pass: the z plan with its smallest radix moved to the front, so the
// first forward pass and the last inverse pass (the reversed plan) run on every thread of the
// block and the real-space work between them stays in registers.
template <int X, class In, class Out> struct DropOne;
template <int X, int... Out> struct DropOne<X, Radices<>, Radices<Out...>> { using type = Radices<Out...>; };
template <int X, int R0, int... Rs, int... Out> struct DropOne<X, Radices<R0, Rs...>, Radices<Out...>> {
  using type = typename std::conditional<R0 == X, Radices<Out..., Rs...>,
                                         typename DropOne<X, Radices<Rs...>, Radices<Out..., R0>>::type>::type;
};
template <int X, class P> struct Prepend;
template <int X, int... Rs> struct Prepend<X, Radices<Rs...>> { using type = Radices<X, Rs...>; };
// Base plan of the middle pass: the z plan, except 3*4*8 at M = 96 (with 12*8 the middle pass
// at 192^3 took 82.5 against 68.7 us, while the other 96-point z passes gain)
template <int M> struct ZMidBase {
  using type = typename ZPlanOf<M>::type;
  static constexpr int rmin = ZPlanOf<M>::rmin;
};
template <> struct ZMidBase<96> {
  using type = Radices<3, 4, 8>;
  static constexpr int rmin = 3;
};
template <int M> struct ZMidPlan {
  static constexpr int r0 = ZMidBase<M>::rmin;
  using fwd = typename Prepend<r0, typename DropOne<r0, typename ZMidBase<M>::type, Radices<>>::type>::type;
  using inv = typename RevImpl<fwd, Radices<>>::type;
};

// Waves-per-SIMD floor of the middle pass: 4 (at most 128 VGPRs) for the shapes that fit
// there without spilling (n >= 256 with three variants), else no floor.
template <int M, int VT> constexpr int zmid_wpe() {
  return GSK_ZMID_WPE > 0 ? GSK_ZMID_WPE : ((M >= 128 && VT == 3) ? 4 : 1);
}

// rows per z tile and the exact block size (also used by the host)
__host__ __device__ constexpr int zrows(int M) {
  return M == 32 ? GSK_ZROWS_32
                 : (M < 32 ? 32
                           : (M == 64 ? GSK_ZROWS_64
                                      : (M == 96 ? GSK_ZROWS_96 : (M < 128 ? 16 : (M == 128 ? GSK_ZROWS_128 : 8)))));
}
__host__ __device__ constexpr int zmid_rows(int M) {
  return M == 256 ? GSK_ZMID_ROWS_256 : (M == 128 ? GSK_ZMID_ROWS_128 : (M == 64 ? GSK_ZMID_ROWS_64 : zrows(M)));
}
// threads of a tile of L rows: one per butterfly of the smallest radix
__host__ __device__ constexpr int zthreads_lr(int L, int M, int rmin) {
  return L * (M / rmin) < 64 ? 64 : (L * (M / rmin) + 63) / 64 * 64;
}
__host__ __device__ constexpr int zthreads_of(int M, int rmin) { return zthreads_lr(zrows(M), M, rmin); }
template <int M> constexpr int zthreads() { return zthreads_of(M, ZPlanOf<M>::rmin); }
template <int M> constexpr int zmid_threads() { return zthreads_lr(zmid_rows(M), M, ZMidBase<M>::rmin); }

// Row pitch of a z tile of L rows in LDS (complex elements): the smallest P >= M + 1 with
// P = 16/L (mod 16) for L <= 16 rows per tile, P odd for L = 32. An FFT pass maps lanes
// line fastest, and the LDS serves an 8-byte access 16 lanes at a time from 16 bank pairs:
// those 16 lanes (L lines x 16/L consecutive positions) then hit 16 distinct pairs.
// (Measured at n = 256, middle pass: P = 130 135 us, 129 141 us, 132 147 us, 136 173 us.)
// GSK_ZPITCH_R overrides the residue mod 16.
__host__ __device__ constexpr int zpitch_l(int M, int L) {
#ifdef GSK_ZPITCH_R
  const int r = GSK_ZPITCH_R;
#else
  const int r = L >= 16 ? 1 : 16 / L;
#endif
  int p = M + 1;
  while (p % 16 != r % 16) ++p;
  return p;
}
__host__ __device__ constexpr int zpitch(int M) { return zpitch_l(M, zrows(M)); }

enum ZOp { Z_COPY = 0, Z_SQUARE = 1, Z_PREC_LOCAL = 2 };

// --------------------------------------------------------------------- split formulas
// C2R: half-spectrum pair (X_k, X_{M-k}) -> packed spectrum pair (Z_k, Z_{M-k}), w = exp(-2 pi i k/n):
//   E' = X_k + conj X_{M-k},  O' = (X_k - conj X_{M-k}) conj(w),  Z_k = E' + i O',  Z_{M-k} = conj(E') + i conj(O')
__device__ __forceinline__ void c2r_pair(float2 xk, float2 xm, float2 w, float2& zk, float2& zm) {
  const float2 E = make_float2(xk.x + xm.x, xk.y - xm.y);
  const float2 D = make_float2(xk.x - xm.x, xk.y + xm.y);
  const float2 O = cmul(D, make_float2(w.x, -w.y));
  zk = make_float2(E.x - O.y, E.y + O.x);
  zm = make_float2(E.x + O.y, -E.y + O.x);
}
// R2C: packed pair (Z_k, Z_{M-k}) -> (X_k, X_{M-k}):
//   E = (Z_k + conj Z_{M-k})/2,  O = (Z_k - conj Z_{M-k})/(2i),  X_k = E + w O,  X_{M-k} = conj(E - w O)
__device__ __forceinline__ void r2c_pair(float2 zk, float2 zm, float2 w, float2& xk, float2& xm) {
  const float2 E = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
  const float2 O = make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
  const float2 WO = cmul(w, O);
  xk = make_float2(E.x + WO.x, E.y + WO.y);
  xm = make_float2(E.x - WO.x, -(E.y - WO.y));
}

// Element `elem` of `base` through a 32-bit byte offset (uniform base + one offset register).
template <class T> __device__ __forceinline__ T& at32(T* base, unsigned elem) {
  using C = typename std::conditional<std::is_const<T>::value, const char, char>::type;
  return *reinterpret_cast<T*>(reinterpret_cast<C*>(base) + (unsigned)(elem * (unsigned)sizeof(T)));
}

// --------------------------------------------------------------------- one-field tile I/O
// TM: work-group size as a multiple of the FFT's thread count (more threads, fewer tasks each);
// LR: rows per tile, P: their LDS row pitch; RM: smallest radix of the tile's plan
template <int M, int TM = 1, int LR = zrows(M), int RM = ZPlanOf<M>::rmin>
struct ZShape {
  static constexpr int L = LR, T = zthreads_lr(LR, M, RM) * TM, H2 = M / 2, P = zpitch_l(M, LR);
  static constexpr int NT = (L * H2 + T - 1) / T;  // tasks per thread
  static constexpr bool FULL = (L * H2) % T == 0;
  __device__ static bool valid(int e) { return FULL || e < L * H2; }
};

// The half-spectrum values of one tile that a thread owns, held in registers: the pair (k, M - k)
// of each of its tasks (k = 0: the Nyquist value kz = M as the partner), and for threads tid < L
// the self-paired value k = M/2 of tile row tid.
template <int M, int TM = 1, int LR = zrows(M), int RM = ZPlanOf<M>::rmin>
struct ZSpec {
  using S = ZShape<M, TM, LR, RM>;
  static_assert(S::T >= S::L, "one thread per row for the self-paired values");
  float2 a[S::NT], b[S::NT], cq;
  // Rows r0.. of one field: a uniform field base and 32-bit element offsets (a field holds
  // fewer than 2^29 elements), so the loads use a scalar base and one offset register each.
  __device__ __forceinline__ void load(const float2* __restrict__ field, const Geo& g, long r0) {
    const unsigned rb = (unsigned)r0 * M, nb = (unsigned)(g.nq + r0);
#pragma unroll
    for (int i = 0; i < S::NT; ++i) {
      const unsigned e = threadIdx.x + i * S::T;
      if (S::valid(e)) {
        const unsigned l = e / S::H2, k = e % S::H2, ro = rb + l * M;
        a[i] = at32(field, ro + k);
        b[i] = at32(field, k == 0 ? nb + l : ro + (M - k));  // k = 0: the kz = M value (Nyquist plane)
      }
    }
    if (threadIdx.x < (unsigned)S::L) cq = at32(field, rb + threadIdx.x * M + S::H2);
  }
  // C2R split into the packed spectra of the LDS tile (row pitch S::P > M). The pair of k = 0,
  // with the imaginary parts of X_0 and X_M taken as zero, is split by the general formula at
  // w = 1 into Z_0 = (X_0 + X_M, X_0 - X_M); its second output lands in the row's padding slot M.
  __device__ __forceinline__ void split(float2* tile, const float2* rt, int tid = threadIdx.x) const {
#pragma unroll
    for (int i = 0; i < S::NT; ++i) {
      const unsigned e = (unsigned)tid + i * S::T;
      if (S::valid(e)) {
        const unsigned l = e / S::H2, k = e % S::H2;
        float2* z = tile + l * S::P;
        const bool k0 = k == 0;
        float2 xk = a[i], xm = b[i];
        xk.y = k0 ? 0.f : xk.y;
        xm.y = k0 ? 0.f : xm.y;
        float2 zk, zm;
        c2r_pair(xk, xm, rt[k], zk, zm);
        z[k] = zk;
        z[M - k] = zm;
      }
    }
    if ((unsigned)tid < (unsigned)S::L) {
      float2 zh, unused;
      c2r_pair(cq, cq, rt[S::H2], zh, unused);
      tile[tid * S::P + S::H2] = zh;
    }
  }
};

// packed spectra of the LDS tile -> half-spectrum rows r0.. of one field (R2C direction);
// offsets as in ZSpec::load. The pair of k = 0 is Z_0 with itself: the general formula at w = 1
// gives X_0 = Z_0.x + Z_0.y and X_M = Z_0.x - Z_0.y (imaginary parts zero); X_{M/2} = conj Z_{M/2}
// of row tid is written by thread tid < L.
template <int M, int TM = 1, int LR = zrows(M), int RM = ZPlanOf<M>::rmin>
__device__ __forceinline__ void z_store_spec(const float2* tile, float2* __restrict__ field, const Geo& g, long r0,
                                             const float2* rt, int tid = threadIdx.x) {
  using S = ZShape<M, TM, LR, RM>;
  const unsigned rb = (unsigned)r0 * M, nb = (unsigned)(g.nq + r0);
#pragma unroll
  for (int i = 0; i < S::NT; ++i) {
    const unsigned e = (unsigned)tid + i * S::T;
    if (S::valid(e)) {
      const unsigned l = e / S::H2, k = e % S::H2, ro = rb + l * M;
      const float2* z = tile + l * S::P;
      const bool k0 = k == 0;
      float2 xk, xm;
      r2c_pair(z[k], z[k0 ? 0 : M - k], rt[k], xk, xm);
      xk.y = k0 ? 0.f : xk.y;
      xm.y = k0 ? 0.f : xm.y;
      at32(field, ro + k) = xk;
      at32(field, k0 ? nb + l : ro + (M - k)) = xm;
    }
  }
  if ((unsigned)tid < (unsigned)S::L) {
    const float2 zh = tile[tid * S::P + S::H2];
    at32(field, rb + (unsigned)tid * M + S::H2) = make_float2(zh.x, -zh.y);
  }
}

// quad (l, q) of the tile: voxels 4q..4q+3 of row l
template <int M, int TM = 1>
__device__ __forceinline__ void z_quad(int i, int& l, int& q) {
  using S = ZShape<M, TM>;
  const int e = threadIdx.x + i * S::T;
  l = e / S::H2;
  q = e % S::H2;
}

template <int M>
__device__ __forceinline__ void z_twiddles(float2* tw, float2* rt, const float2* __restrict__ twM,
                                           const float2* __restrict__ rtn) {
  for (int i = threadIdx.x; i < M; i += blockDim.x) {
    tw[i] = twM[i];
    rt[i] = rtn[i];
  }
  __syncthreads();
}

template <int M, bool FWD>
__device__ __forceinline__ void z_fft1(float2* tile, const float2* tw) {
  constexpr int L = zrows(M);
  const Lines Ls{L, L, zpitch(M), 0, 1};
  fft_lines_p<float2, M, FWD, typename ZPlanOf<M>::type>(tile, Ls, tw);
}

// --------------------------------------------------------------------- forward z pass (one field per block)
// real field f = blockIdx.y -> (square) -> R2C -> half-spectrum rows.
// OP = Z_SQUARE: eta_p -> eta_p^2;  OP = Z_COPY: unchanged.
template <int M, int OP>
__global__ void __launch_bounds__(zthreads<M>()) k_zfwd(const float* __restrict__ in, float2* __restrict__ kout, Geo g,
                                                       const float2* __restrict__ twM, const float2* __restrict__ rtn) {
  using S = ZShape<M>;
  constexpr int n = 2 * M;
  extern __shared__ float2 smem[];
  float2* tw = smem;
  float2* rt = smem + M;
  float2* tile = smem + 2 * M;
  const int f = geo_field(g);
  const long r0 = (long)geo_tile(g) * S::L;
  const float* src = in + f * g.nvox + r0 * n;
  float4 v[S::NT];
#pragma unroll
  for (int i = 0; i < S::NT; ++i) {
    int l, q;
    z_quad<M>(i, l, q);
    if (S::valid(threadIdx.x + i * S::T)) v[i] = *reinterpret_cast<const float4*>(src + l * n + 4 * q);
  }
  z_twiddles<M>(tw, rt, twM, rtn);
#pragma unroll
  for (int i = 0; i < S::NT; ++i) {
    int l, q;
    z_quad<M>(i, l, q);
    if (S::valid(threadIdx.x + i * S::T)) {
      float4 t = v[i];
      if (OP == Z_SQUARE) t = make_float4(t.x * t.x, t.y * t.y, t.z * t.z, t.w * t.w);
      float2* row = tile + l * zpitch(M);
      row[2 * q] = make_float2(t.x, t.y);
      row[2 * q + 1] = make_float2(t.z, t.w);
    }
  }
  __syncthreads();
  z_fft1<M, true>(tile, tw);
  z_store_spec<M>(tile, kout + f * g.kfs, g, r0, rt);
}

// --------------------------------------------------------------------- inverse z pass (one field per block)
// half-spectrum rows of field f = blockIdx.y -> C2R -> real field.
// adv (may be null): step counter advanced once by the launch (the last kernel of a fused step).
template <int M>
__global__ void __launch_bounds__(zthreads<M>()) k_zinv(const float2* __restrict__ kin, float* __restrict__ out, Geo g,
                                                       const float2* __restrict__ twM, const float2* __restrict__ rtn,
                                                       uint32_t* adv) {
  if (adv && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *adv += 1u;
  using S = ZShape<M>;
  constexpr int n = 2 * M;
  extern __shared__ float2 smem[];
  float2* tw = smem;
  float2* rt = smem + M;
  float2* tile = smem + 2 * M;
  const int f = geo_field(g);
  const long r0 = (long)geo_tile(g) * S::L;
  ZSpec<M> sp;
  sp.load(kin + f * g.kfs, g, r0);
  z_twiddles<M>(tw, rt, twM, rtn);
  sp.split(tile, rt);
  __syncthreads();
  z_fft1<M, false>(tile, tw);
  float* dst = out + f * g.nvox + r0 * n;
#pragma unroll
  for (int i = 0; i < S::NT; ++i) {
    int l, q;
    z_quad<M>(i, l, q);
    if (S::valid(threadIdx.x + i * S::T)) {
      const float2* row = tile + l * zpitch(M);
      const float2 a = row[2 * q], b = row[2 * q + 1];
      *reinterpret_cast<float4*>(dst + l * n + 4 * q) = make_float4(a.x, a.y, b.x, b.y);
    }
  }
}

// --------------------------------------------------------------------- inverse z pass + next forward square
// The last kernel of a martensite step fused with the first kernel of the next one: half-spectrum
// rows of eta_f -> C2R -> eta_f written to the real field -> eta_f^2 -> R2C -> half-spectrum rows
// written back in place (the z-transformed S_f = eta_f^2 the next step starts from). The squares
// are taken of the FP32 values written to the real field, so S_f is bit-identical to a separate
// forward pass over that field.
// adv (may be null): step counter advanced once by the launch.
template <int M, bool BT = false>
__global__ void __launch_bounds__(zthreads<M>()) k_zinv_sq(float2* __restrict__ kbuf, float* __restrict__ out, Geo g,
                                                          const float2* __restrict__ twM,
                                                          const float2* __restrict__ rtn, uint32_t* adv,
                                                          const MemberPtrs* __restrict__ mt) {
  if (BT && mt) {  // batched ensemble launch
    kbuf = mt[blockIdx.z].kbuf;
    out = mt[blockIdx.z].fld;
    adv = mt[blockIdx.z].dstep;
  }
  if (adv && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *adv += 1u;
  using S = ZShape<M>;
  constexpr int n = 2 * M;
  extern __shared__ float2 smem[];
  float2* tw = smem;
  float2* rt = smem + M;
  float2* tile = smem + 2 * M;
  const int f = geo_field(g);
  const long r0 = (long)geo_tile(g) * S::L;
  float2* field = kbuf + f * g.kfs;
  {
    ZSpec<M> sp;
    sp.load(field, g, r0);
    z_twiddles<M>(tw, rt, twM, rtn);
    sp.split(tile, rt);
  }
  __syncthreads();
  z_fft1<M, false>(tile, tw);
  float* dst = out + f * g.nvox + r0 * n;
#pragma unroll
  for (int i = 0; i < S::NT; ++i) {
    int l, q;
    z_quad<M>(i, l, q);
    if (S::valid(threadIdx.x + i * S::T)) {
      float2* row = tile + l * zpitch(M);
      const float2 a = row[2 * q], b = row[2 * q + 1];
      at32(reinterpret_cast<float4*>(dst), (unsigned)(l * n + 4 * q) / 4) = make_float4(a.x, a.y, b.x, b.y);
      // each thread squares the very elements it has just read: no barrier between
      row[2 * q] = make_float2(a.x * a.x, a.y * a.y);
      row[2 * q + 1] = make_float2(b.x * b.x, b.y * b.y);
    }
  }
  __syncthreads();
  z_fft1<M, true>(tile, tw);
  z_store_spec<M>(tile, field, g, r0, rt);
}

// --------------------------------------------------------------------- martensite middle z pass
// For each variant p in turn: R_hat_p rows -> C2R -> F_p = f'(eta) + 2 w eta_p R_p + a xi (Philox in
// registers) -> R2C, in place. The last inverse Stockham pass leaves thread (l, j) = (b % L, b / L)
// holding packed positions m = j + r*M/R0 of row l, i.e. the voxel pairs (2m, 2m+1); the force is
// formed on those registers and the first forward pass starts from them. VT > 0: the variant count
// is a compile-time constant and every eta_q the thread needs is prefetched into registers;
// VT = 0: any variant count, eta read when needed.
template <int M, int VT, bool BT = false>
__global__ void __launch_bounds__(zmid_threads<M>()) __attribute__((amdgpu_waves_per_eu(zmid_wpe<M, VT>())))
k_zmid_mart(float2* __restrict__ kbuf, const float* __restrict__ eta, Geo g, MartK mk, const float2* __restrict__ twM,
            const float2* __restrict__ rtn, const MemberPtrs* __restrict__ mt) {
  if (BT && mt) {  // batched ensemble launch
    kbuf = mt[blockIdx.z].kbuf;
    eta = mt[blockIdx.z].fld;
    mk = mt[blockIdx.z].mk;
  }
  constexpr int LZ = zmid_rows(M), RMZ = ZMidBase<M>::rmin;
  using S = ZShape<M, 1, LZ, RMZ>;
  using P = ZMidPlan<M>;
  constexpr int n = 2 * M, R0 = P::r0, NB = S::L * (M / R0);
  constexpr int VR = VT > 0 ? VT : 1;
  extern __shared__ float2 smem[];
  float2* tw = smem;
  float2* rt = smem + M;
  float2* tile = smem + 2 * M;
  const long r0 = (long)blockIdx.x * S::L;
  const int F = VT > 0 ? VT : mk.F;
  const int b = threadIdx.x;
  const bool act = b < NB;
  const int l = b % S::L, j = b / S::L;
  // voxel of register r: row r0 + l, pair m = j + r*M/R0
  const float* erow = eta + r0 * n;        // uniform base of the tile's rows
  const unsigned vloc = l * n + 2 * j;      // voxel of register 0 relative to erow
  float2 pre[R0][VR];
  if constexpr (VT > 0) {
    if (act) {
#pragma unroll
      for (int r = 0; r < R0; ++r)
#pragma unroll
        for (int f = 0; f < VR; ++f)
          pre[r][f] = at32(reinterpret_cast<const float2*>(erow + f * g.nvox), vloc / 2 + r * (M / R0));
    }
  }
  ZSpec<M, 1, LZ, RMZ> sp;
  sp.load(kbuf, g, r0);
  z_twiddles<M>(tw, rt, twM, rtn);
  const Lines Ls{S::L, S::L, S::P, 0, 1};
  auto no_ld = [](int, int) { return float2{}; };
  auto no_st = [](int, int, float2) {};
  const uint32_t step = step_of(mk.step, mk.stepp);
#pragma unroll 1
  for (int p = 0; p < F; ++p) {
    // an opaque copy of the thread index per variant: the LDS addresses of the passes are
    // recomputed in each iteration instead of being held in registers across the loop
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    sp.split(tile, rt, tid);
    __syncthreads();
    if (p + 1 < F) sp.load(kbuf + (p + 1) * g.kfs, g, r0);
    float2 v[R0];
    gchain<float2, M, false, 0, 2, 1>(tile, Ls, tw, no_ld, no_st, v, typename P::inv{}, tid);
    if (act) {
      // Noise of the thread's voxel pairs. A Philox block covers 4 voxels, i.e. the pair of this
      // thread and that of lane tid ^ L (j of opposite parity, same row). With ND each lane
      // computes half of the blocks the two share and hands the partner the half it needs,
      // so every block is computed once; bit-identical to normal2 for each pair.
      constexpr bool ND = R0 % 2 == 0 && (M / R0) % 2 == 0 && S::L <= 32;
      float xin[R0][2];
      if constexpr (ND) {
        if (mk.noise != 0.f) {
          const uint32_t h = (uint32_t)(tid / S::L) & 1u;
#pragma unroll
          for (int q = 0; q < R0 / 2; ++q) {
            const unsigned ro = h * (R0 / 2) + q;
            uint32_t gq = ((unsigned)r0 * n + vloc + 2 * ro * (M / R0)) >> 2;
            asm volatile("" : "+v"(gq));  // opaque per variant (see below)
            uint32_t c[4] = {gq, (uint32_t)p, step, 0u};
            philox10(c, (uint32_t)mk.seed, (uint32_t)(mk.seed >> 32));
            const uint32_t a0 = h ? c[2] : c[0], a1 = h ? c[3] : c[1];
            const uint32_t b0 = __shfl_xor(h ? c[0] : c[2], S::L), b1 = __shfl_xor(h ? c[1] : c[3], S::L);
            float own[2], oth[2];
            box_muller(a0, a1, own[0], own[1]);
            box_muller(b0, b1, oth[0], oth[1]);
            xin[q][0] = h ? oth[0] : own[0];
            xin[q][1] = h ? oth[1] : own[1];
            xin[R0 / 2 + q][0] = h ? own[0] : oth[0];
            xin[R0 / 2 + q][1] = h ? own[1] : oth[1];
          }
        } else {
#pragma unroll
          for (int r = 0; r < R0; ++r) xin[r][0] = xin[r][1] = 0.f;
        }
      }
#pragma unroll
      for (int r = 0; r < R0; ++r) {
        const unsigned vo = vloc + 2 * r * (M / R0);
        const unsigned vox = (unsigned)r0 * n + vo;  // < n^3 <= 2^27
        float e0 = 0.f, e1 = 0.f, s0 = 0.f, s1 = 0.f;
        if constexpr (VT > 0) {
#pragma unroll
          for (int f = 0; f < VR; ++f) {
            const float2 t = pre[r][f];
            s0 += t.x * t.x; s1 += t.y * t.y;
            if (f == p) { e0 = t.x; e1 = t.y; }
          }
        } else {
          for (int f = 0; f < F; ++f) {
            const float2 t = at32(reinterpret_cast<const float2*>(erow + f * g.nvox), vo / 2);
            s0 += t.x * t.x; s1 += t.y * t.y;
            if (f == p) { e0 = t.x; e1 = t.y; }
          }
        }
        float xi[2] = {0.f, 0.f};
        if constexpr (ND) {
          xi[0] = xin[r][0];
          xi[1] = xin[r][1];
        } else if (mk.noise != 0.f) {
          // opaque per variant: keeps the compiler from hoisting the variant-independent part
          // of the Philox rounds out of the variant loop (it would hold it in registers)
          uint32_t gq = vox >> 2;
          asm volatile("" : "+v"(gq));
          normal2(mk.seed, gq, (uint32_t)p, step, (vox >> 1) & 1u, xi);
        }
        const float o0 = mk.A * e0 - mk.B * e0 * e0 + mk.C * e0 * s0 + mk.two_wel * e0 * v[r].x + mk.noise * xi[0];
        const float o1 = mk.A * e1 - mk.B * e1 * e1 + mk.C * e1 * s1 + mk.two_wel * e1 * v[r].y + mk.noise * xi[1];
        v[r] = make_float2(o0, o1);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    gchain<float2, M, true, 2, 0, 1>(tile, Ls, tw, no_ld, no_st, v, typename P::fwd{}, tid);
    z_store_spec<M, 1, LZ, RMZ>(tile, kbuf + p * g.kfs, g, r0, rt, tid);
    __syncthreads();  // the next variant's split overwrites the tile
  }
}

// --------------------------------------------------------------------- precipitate forward z pass
// All fields of a tile in LDS: (dc, eta_1..V) -> (mu_c, F_1..F_V) with noise on eta, then R2C.
// R (may be null): real elastic fields R_v of the eta^2 source, adding el * R_v eta_v to F_v;
// cnoise: amplitude of noise on mu_c (Philox field 0x8000), both used by the reference rule.
template <int M>
__global__ void __launch_bounds__(zthreads<M>()) k_zfwd_local(const float* __restrict__ in, float2* __restrict__ kout,
                                                             Geo g, PrecK pk, const float2* __restrict__ twM,
                                                             const float2* __restrict__ rtn,
                                                             const float* __restrict__ R, float el, float cnoise) {
  using S = ZShape<M>;
  constexpr int L = S::L, Q = M / 2, n = 2 * M;
  extern __shared__ float2 smem[];
  float2* tw = smem;
  float2* rt = smem + M;
  float2* tile = smem + 2 * M;
  const long r0 = (long)blockIdx.x * L;
  const int F = pk.V + 1;
  z_twiddles<M>(tw, rt, twM, rtn);
  for (int e = threadIdx.x; e < L * Q; e += blockDim.x) {
    const int l = e / Q, q = e % Q;
    const long vox = (r0 + l) * n + 4 * q;
    float v[8][4];
    for (int f = 0; f < F; ++f) {
      const float4 t = *reinterpret_cast<const float4*>(in + f * g.nvox + vox);
      v[f][0] = t.x; v[f][1] = t.y; v[f][2] = t.z; v[f][3] = t.w;
    }
    const int V = pk.V;
    float out[8][4];
    float xc[4] = {0.f, 0.f, 0.f, 0.f};
    if (cnoise != 0.f) normal4(pk.seed, (uint32_t)(vox >> 2), 0x8000u, step_of(pk.step, pk.stepp), xc);
    for (int i = 0; i < 4; ++i) {
      const float c = pk.cbar + v[0][i];
      float s2 = 0.f;
      for (int a = 1; a <= V; ++a) s2 += v[a][i] * v[a][i];
      out[0][i] = 0.5f * pk.A2 * s2 + pk.A1 * (c - pk.c1) + cnoise * xc[i];
    }
    for (int a = 1; a <= V; ++a) {
      float xi[4] = {0.f, 0.f, 0.f, 0.f};
      if (pk.noise != 0.f) normal4(pk.seed, (uint32_t)(vox >> 2), (uint32_t)(a - 1), step_of(pk.step, pk.stepp), xi);
      for (int i = 0; i < 4; ++i) {
        const float et = v[a][i];
        float t2 = 0.f, t4 = 0.f;
        for (int b = 1; b <= V; ++b) {
          if (b == a) continue;
          const float qq = v[b][i] * v[b][i];
          t2 += qq;
          t4 += qq * qq;
        }
        // ordered pairs i != j, both != a: (sum q)^2 - sum q^2
        const float t22 = t2 * t2 - t4;
        const float c = pk.cbar + v[0][i];
        const float e2 = et * et;
        out[a][i] = (2.f * pk.A5 * et + 4.f * pk.A6 * e2 * et) * t2 + 2.f * pk.A6 * et * t4 +
                    2.f * pk.A7 * et * t22 + pk.A2 * et * (c - pk.c2) - pk.A3 * e2 * et + pk.A4 * e2 * e2 * et +
                    pk.noise * xi[i];
        if (R) out[a][i] += el * R[(a - 1) * g.nvox + vox + i] * et;
      }
    }
    for (int f = 0; f < F; ++f) {
      float2* row = tile + (f * L + l) * zpitch(M);
      row[2 * q] = make_float2(out[f][0], out[f][1]);
      row[2 * q + 1] = make_float2(out[f][2], out[f][3]);
    }
  }
  __syncthreads();
  for (int f = 0; f < F; ++f) z_fft1<M, true>(tile + f * L * zpitch(M), tw);
  for (int f = 0; f < F; ++f) z_store_spec<M>(tile + f * L * zpitch(M), kout + f * g.kfs, g, r0, rt);
}

// Local precipitate force of output field f at four voxels from the FT = V + 1 field values
// e[u][k] (u = 0: the concentration fluctuation about cbar, u = 1..V: the order parameters):
//   f = 0: mu_c = A2/2 sum_a eta_a^2 + A1 (c - c1) + cnoise xi
//   f > 0: F_f = (2 A5 eta + 4 A6 eta^3) t2 + 2 A6 eta t4 + 2 A7 eta t22 + A2 eta (c - c2) - A3 eta^3
//                + A4 eta^5 + noise xi [+ el R eta]
// with t2, t4 the sums of eta_b^2, eta_b^4 over b != f and t22 = t2^2 - t4 (ordered pairs of other
// variants). xi: the normal deviates of field f's noise block; rk (hasR): the elastic R of eta_f.
template <int FT>
__device__ __forceinline__ void prec_local4(const PrecK& pk, int f, const float (&e)[FT][4], const float (&xi)[4],
                                            float cnoise, bool hasR, const float (&rk)[4], float el, float (&out)[4]) {
  constexpr int V = FT - 1;
  if (f == 0) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      float s2 = 0.f;
#pragma unroll
      for (int a = 1; a <= V; ++a) s2 += e[a][k] * e[a][k];
      out[k] = 0.5f * pk.A2 * s2 + pk.A1 * (pk.cbar + e[0][k] - pk.c1) + cnoise * xi[k];
    }
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      float et = 0.f, t2 = 0.f, t4 = 0.f;
#pragma unroll
      for (int b = 1; b <= V; ++b) {
        const float x = e[b][k];
        if (b == f) {
          et = x;
        } else {
          const float qq = x * x;
          t2 += qq;
          t4 += qq * qq;
        }
      }
      const float t22 = t2 * t2 - t4;
      const float c = pk.cbar + e[0][k];
      const float e2 = et * et;
      float o = (2.f * pk.A5 * et + 4.f * pk.A6 * e2 * et) * t2 + 2.f * pk.A6 * et * t4 + 2.f * pk.A7 * et * t22 +
                pk.A2 * et * (c - pk.c2) - pk.A3 * e2 * et + pk.A4 * e2 * e2 * et + pk.noise * xi[k];
      if (hasR) o += el * rk[k] * et;
      out[k] = o;
    }
  }
}

// Waves-per-SIMD floor of the register-resident precipitate z passes: 4 (at most 128 VGPRs)
// for n >= 256.
template <int M, int FT> constexpr int zlocal_wpe() { return M >= 128 ? 4 : 1; }

// Local precipitate forces of one tile from its FT = V + 1 fields held in registers (v[i][f]:
// quad i of the thread, field f), then the R2C of each output field over a single LDS tile.
// Inside the field loop the thread index and the Philox counter are made opaque, so the
// compiler recomputes the LDS addresses and the field-independent Philox work per field
// instead of holding them in registers.
template <int M, int FT>
__device__ __forceinline__ void zlocal_forces_r(const float4 (&v)[ZShape<M>::NT][FT], float2* tile, const float2* tw,
                                                const float2* rt, float2* __restrict__ kout, const Geo& g,
                                                const PrecK& pk, long r0, const float* __restrict__ R, float el,
                                                float cnoise) {
  using S = ZShape<M>;
  constexpr int n = 2 * M;
  const Lines Ls{S::L, S::L, zpitch(M), 0, 1};
  auto no_ld = [](int, int) { return float2{}; };
  auto no_st = [](int, int, float2) {};
  const uint32_t step = step_of(pk.step, pk.stepp);
#pragma unroll 1
  for (int f = 0; f < FT; ++f) {
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
#pragma unroll
    for (int i = 0; i < S::NT; ++i) {
      const int e0 = tid + i * S::T;
      if (!S::valid(e0)) continue;
      const int l = e0 / S::H2, q = e0 % S::H2;
      const unsigned vox = (unsigned)r0 * n + l * n + 4 * q;  // < n^3 <= 2^27
      uint32_t gq = vox >> 2;
      asm volatile("" : "+v"(gq));
      float e[FT][4];
#pragma unroll
      for (int u = 0; u < FT; ++u) {
        e[u][0] = v[i][u].x; e[u][1] = v[i][u].y; e[u][2] = v[i][u].z; e[u][3] = v[i][u].w;
      }
      float out[4];
      float xi[4] = {0.f, 0.f, 0.f, 0.f};
      float rk[4] = {0.f, 0.f, 0.f, 0.f};
      if (f == 0) {
        if (cnoise != 0.f) normal4(pk.seed, gq, 0x8000u, step, xi);
      } else {
        if (pk.noise != 0.f) normal4(pk.seed, gq, (uint32_t)(f - 1), step, xi);
        if (R) {
          const float4 rq =
              at32(reinterpret_cast<const float4*>(R + (f - 1) * g.nvox + r0 * n), (unsigned)(l * n + 4 * q) / 4);
          rk[0] = rq.x; rk[1] = rq.y; rk[2] = rq.z; rk[3] = rq.w;
        }
      }
      prec_local4<FT>(pk, f, e, xi, cnoise, R != nullptr, rk, el, out);
      float2* row = tile + l * zpitch(M);
      row[2 * q] = make_float2(out[0], out[1]);
      row[2 * q + 1] = make_float2(out[2], out[3]);
    }
    __syncthreads();
    gchain<float2, M, true, 0, 0, 1>(tile, Ls, tw, no_ld, no_st, (float2*)nullptr, typename ZPlanOf<M>::type{}, tid);
    z_store_spec<M>(tile, kout + f * g.kfs, g, r0, rt, tid);
    __syncthreads();  // the next field overwrites the tile
  }
}

// k_zfwd_local for a compile-time field count FT = V + 1 (2..5): every field of the tile is
// loaded into registers first (16-byte quads), then the output fields are built and
// transformed one after the other over a single LDS tile.
// Waves-per-SIMD floor 4 (at most 128 VGPRs) for n >= 256.
template <int M, int FT>
__global__ void __launch_bounds__(zthreads<M>()) __attribute__((amdgpu_waves_per_eu(zlocal_wpe<M, FT>())))
k_zfwd_local_r(const float* __restrict__ in, float2* __restrict__ kout, Geo g, PrecK pk,
               const float2* __restrict__ twM, const float2* __restrict__ rtn, const float* __restrict__ R, float el,
               float cnoise) {
  using S = ZShape<M>;
  constexpr int n = 2 * M;
  extern __shared__ float2 smem[];
  float2* tw = smem;
  float2* rt = smem + M;
  float2* tile = smem + 2 * M;
  const long r0 = (long)blockIdx.x * S::L;
  const float* src = in + r0 * n;
  float4 v[S::NT][FT];
#pragma unroll
  for (int i = 0; i < S::NT; ++i) {
    int l, q;
    z_quad<M>(i, l, q);
    if (S::valid(threadIdx.x + i * S::T)) {
#pragma unroll
      for (int f = 0; f < FT; ++f)
        v[i][f] = at32(reinterpret_cast<const float4*>(src + f * g.nvox), (unsigned)(l * n + 4 * q) / 4);
    }
  }
  z_twiddles<M>(tw, rt, twM, rtn);
  zlocal_forces_r<M, FT>(v, tile, tw, rt, kout, g, pk, r0, R, el, cnoise);
}

// The last kernel of a precipitate step fused with the first kernel of the next one: for each
// field in turn, half-spectrum rows -> C2R -> the real field (written out and kept in registers;
// the next field's spectrum is loaded during the transform); then the local forces of the next
// step from the registers, R2C, written back in place. The forces are formed from the FP32
// values written to the real fields, so the result is bit-identical to k_zfwd_local_r over them.
// The noise counter is read as the next step's (the step's x update has advanced it).
template <int M, int FT>
__global__ void __launch_bounds__(zthreads<M>()) __attribute__((amdgpu_waves_per_eu(zlocal_wpe<M, FT>())))
k_zinv_local_r(float2* __restrict__ kbuf, float* __restrict__ out, Geo g, PrecK pk, const float2* __restrict__ twM,
               const float2* __restrict__ rtn) {
  using S = ZShape<M>;
  constexpr int n = 2 * M;
  extern __shared__ float2 smem[];
  float2* tw = smem;
  float2* rt = smem + M;
  float2* tile = smem + 2 * M;
  const long r0 = (long)blockIdx.x * S::L;
  float* dst = out + r0 * n;
  float4 v[S::NT][FT];
  ZSpec<M> sp;
  sp.load(kbuf, g, r0);
  z_twiddles<M>(tw, rt, twM, rtn);
#pragma unroll
  for (int f = 0; f < FT; ++f) {
    sp.split(tile, rt);
    __syncthreads();
    if (f + 1 < FT) sp.load(kbuf + (f + 1) * g.kfs, g, r0);
    z_fft1<M, false>(tile, tw);
#pragma unroll
    for (int i = 0; i < S::NT; ++i) {
      int l, q;
      z_quad<M>(i, l, q);
      if (S::valid(threadIdx.x + i * S::T)) {
        const float2* row = tile + l * zpitch(M);
        const float2 a = row[2 * q], b = row[2 * q + 1];
        v[i][f] = make_float4(a.x, a.y, b.x, b.y);
        at32(reinterpret_cast<float4*>(dst + f * g.nvox), (unsigned)(l * n + 4 * q) / 4) = v[i][f];
      }
    }
    __syncthreads();  // the next field's split overwrites the tile
  }
  zlocal_forces_r<M, FT>(v, tile, tw, rt, kbuf, g, pk, r0, nullptr, 0.f, 0.f);
}

}  // namespace gsk
