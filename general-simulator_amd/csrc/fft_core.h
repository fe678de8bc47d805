// LDS-staged Stockham FFT engine for CDNA4 (gfx950).
//
// A tile of independent lines lives in LDS. Line l starts at
//   base(l) = (l / lw) * ls1 + (l % lw) * ls0
// and its element p sits at base(l) + p * ps (complex units). One radix-R
// pass maps butterfly b (line-fastest: l = b % nl, j = b / nl) to
//   read  in[j + r*N/R]            r = 0..R-1
//   twiddle by w^(r*k*N/(Ns*R)),   k = j mod Ns, w = exp(-+2 pi i / N)
//   R-point DFT in registers
//   write out[(j/Ns)*Ns*R + k + r*Ns]
// which is the self-sorting Stockham formulation: after the last pass the
// line is in natural order. Each pass reads all operands into registers,
// synchronises, then writes back into the same LDS buffer.
// The block must provide at least nl*N/Rmin threads.
#pragma once
#include <hip/hip_runtime.h>

namespace gsk {

template <typename R> struct V2;
template <> struct V2<float> { using T = float2; };
template <> struct V2<double> { using T = double2; };

template <typename T2> __device__ __forceinline__ T2 mk(decltype(T2::x) a, decltype(T2::x) b) {
  T2 r; r.x = a; r.y = b; return r;
}
template <typename T2> __device__ __forceinline__ T2 cadd(T2 a, T2 b) { return mk<T2>(a.x + b.x, a.y + b.y); }
template <typename T2> __device__ __forceinline__ T2 csub(T2 a, T2 b) { return mk<T2>(a.x - b.x, a.y - b.y); }
template <typename T2> __device__ __forceinline__ T2 cmul(T2 a, T2 b) {
  return mk<T2>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T2> __device__ __forceinline__ T2 cconj(T2 a) { return mk<T2>(a.x, -a.y); }
template <typename T2, typename S> __device__ __forceinline__ T2 cscale(T2 a, S s) { return mk<T2>(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <bool FWD, typename T2> __device__ __forceinline__ T2 rot(T2 a) {
  return FWD ? mk<T2>(a.y, -a.x) : mk<T2>(-a.y, a.x);
}

// ----------------------------------------------------------------- small DFTs
template <bool FWD, typename T2> __device__ __forceinline__ void dft2(T2* v) {
  T2 a = v[0], b = v[1];
  v[0] = cadd(a, b); v[1] = csub(a, b);
}

template <bool FWD, typename T2> __device__ __forceinline__ void dft3(T2* v) {
  using S = decltype(T2::x);
  const S h = S(0.86602540378443864676);  // sqrt(3)/2
  T2 s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
  T2 m = mk<T2>(v[0].x - S(0.5) * s.x, v[0].y - S(0.5) * s.y);
  T2 t = rot<FWD>(d);  // -+i (a1 - a2)
  v[0] = cadd(v[0], s);
  v[1] = mk<T2>(m.x + h * t.x, m.y + h * t.y);
  v[2] = mk<T2>(m.x - h * t.x, m.y - h * t.y);
}

template <bool FWD, typename T2> __device__ __forceinline__ void dft4v(T2& a0, T2& a1, T2& a2, T2& a3) {
  T2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = rot<FWD>(csub(a1, a3));
  a0 = cadd(t0, t2); a2 = csub(t0, t2); a1 = cadd(t1, t3); a3 = csub(t1, t3);
}
template <bool FWD, typename T2> __device__ __forceinline__ void dft4(T2* v) { dft4v<FWD>(v[0], v[1], v[2], v[3]); }

template <bool FWD, typename T2> __device__ __forceinline__ void dft8(T2* v) {
  using S = decltype(T2::x);
  const S c = S(0.70710678118654752440);
  T2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  T2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4v<FWD>(e0, e1, e2, e3);
  dft4v<FWD>(o0, o1, o2, o3);
  // w^1 = (c, -+c), w^2 = -+i, w^3 = (-c, -+c)
  T2 w1o1 = FWD ? mk<T2>(c * (o1.x + o1.y), c * (o1.y - o1.x)) : mk<T2>(c * (o1.x - o1.y), c * (o1.y + o1.x));
  T2 w2o2 = rot<FWD>(o2);
  T2 w3o3 = FWD ? mk<T2>(c * (-o3.x + o3.y), c * (-o3.y - o3.x)) : mk<T2>(c * (-o3.x - o3.y), c * (-o3.y + o3.x));
  v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
  v[1] = cadd(e1, w1o1); v[5] = csub(e1, w1o1);
  v[2] = cadd(e2, w2o2); v[6] = csub(e2, w2o2);
  v[3] = cadd(e3, w3o3); v[7] = csub(e3, w3o3);
}

template <bool FWD, typename T2> __device__ __forceinline__ void dft16(T2* v) {
  using S = decltype(T2::x);
  T2 e[8], o[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) { e[i] = v[2 * i]; o[i] = v[2 * i + 1]; }
  dft8<FWD>(e);
  dft8<FWD>(o);
  // cos(pi k/8), sin(pi k/8) for k = 0..7
  const S cs[8] = {S(1.0), S(0.92387953251128675613), S(0.70710678118654752440), S(0.38268343236508977173),
                   S(0.0), S(-0.38268343236508977173), S(-0.70710678118654752440), S(-0.92387953251128675613)};
  const S sn[8] = {S(0.0), S(0.38268343236508977173), S(0.70710678118654752440), S(0.92387953251128675613),
                   S(1.0), S(0.92387953251128675613), S(0.70710678118654752440), S(0.38268343236508977173)};
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    T2 w = mk<T2>(cs[k], FWD ? -sn[k] : sn[k]);
    T2 t = (k == 0) ? o[k] : cmul(o[k], w);
    v[k] = cadd(e[k], t);
    v[k + 8] = csub(e[k], t);
  }
}

// 12 = 3 x 4: 4-point DFTs over n1 of x[3 n1 + n2], twiddles w12^(n2 k1), 3-point DFTs over n2
// giving X[k1 + 4 k2]
template <bool FWD, typename T2> __device__ __forceinline__ void dft12(T2* v) {
  using S = decltype(T2::x);
  const S c = S(0.86602540378443864676), h = S(0.5);  // cos(pi/6), sin(pi/6)
  T2 a[3][4];
#pragma unroll
  for (int n2 = 0; n2 < 3; ++n2) {
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1) a[n2][n1] = v[3 * n1 + n2];
    dft4v<FWD>(a[n2][0], a[n2][1], a[n2][2], a[n2][3]);
  }
  const S sg = FWD ? S(-1) : S(1);  // w12^m = (cos(pi m/6), sg sin(pi m/6))
  a[1][1] = cmul(a[1][1], mk<T2>(c, sg * h));
  a[1][2] = cmul(a[1][2], mk<T2>(h, sg * c));
  a[1][3] = rot<FWD>(a[1][3]);
  a[2][1] = cmul(a[2][1], mk<T2>(h, sg * c));
  a[2][2] = cmul(a[2][2], mk<T2>(-h, sg * c));
  a[2][3] = mk<T2>(-a[2][3].x, -a[2][3].y);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    T2 t[3] = {a[0][k1], a[1][k1], a[2][k1]};
    dft3<FWD>(t);
    v[k1] = t[0];
    v[k1 + 4] = t[1];
    v[k1 + 8] = t[2];
  }
}

template <int R, bool FWD, typename T2> __device__ __forceinline__ void dftR(T2* v) {
  if constexpr (R == 2) dft2<FWD>(v);
  else if constexpr (R == 3) dft3<FWD>(v);
  else if constexpr (R == 4) dft4<FWD>(v);
  else if constexpr (R == 8) dft8<FWD>(v);
  else if constexpr (R == 12) dft12<FWD>(v);
  else if constexpr (R == 16) dft16<FWD>(v);
  else static_assert(R == 2, "unsupported radix");
}

// ----------------------------------------------------------------- plans
template <int... Rs> struct Radices {};
template <int N> struct PlanOf;
template <> struct PlanOf<2> { using type = Radices<2>; static constexpr int rmin = 2; };
template <> struct PlanOf<4> { using type = Radices<4>; static constexpr int rmin = 4; };
template <> struct PlanOf<8> { using type = Radices<8>; static constexpr int rmin = 8; };
template <> struct PlanOf<16> { using type = Radices<16>; static constexpr int rmin = 16; };
template <> struct PlanOf<32> { using type = Radices<8, 4>; static constexpr int rmin = 4; };
template <> struct PlanOf<48> { using type = Radices<3, 16>; static constexpr int rmin = 3; };
template <> struct PlanOf<64> { using type = Radices<8, 8>; static constexpr int rmin = 8; };
// GSK_PLAN_96: 0 = 3*4*8, 1 = 12*8, 2 = 8*12 (also the z transforms of n = 192, except the
// middle martensite pass, which keeps 3*4*8). 12*8: 96^3 step 0.1019 -> 0.0967 ms, last z pass
// at 192^3 53 -> 46 us; 8*12: 0.0980 ms.
#ifndef GSK_PLAN_96
#define GSK_PLAN_96 1
#endif
#if GSK_PLAN_96 == 1
template <> struct PlanOf<96> { using type = Radices<12, 8>; static constexpr int rmin = 8; };
#elif GSK_PLAN_96 == 2
template <> struct PlanOf<96> { using type = Radices<8, 12>; static constexpr int rmin = 8; };
#else
template <> struct PlanOf<96> { using type = Radices<3, 4, 8>; static constexpr int rmin = 3; };
#endif
template <> struct PlanOf<128> { using type = Radices<16, 8>; static constexpr int rmin = 8; };
// GSK_PLAN_192: 0 = 3*8*8, 1 = 12*16, 2 = 16*12. 12*16 takes two passes and 16 butterflies per
// line like 256 = 16*16 (192^3 step 0.384 -> 0.347 ms against 3*8*8, whose radix-3 pass needs
// 64 butterflies per line and leaves most threads idle in the radix-8 passes; 16*12: 0.350)
#ifndef GSK_PLAN_192
#define GSK_PLAN_192 1
#endif
#if GSK_PLAN_192 == 1
template <> struct PlanOf<192> { using type = Radices<12, 16>; static constexpr int rmin = 12; };
#elif GSK_PLAN_192 == 2
template <> struct PlanOf<192> { using type = Radices<16, 12>; static constexpr int rmin = 12; };
#else
template <> struct PlanOf<192> { using type = Radices<3, 8, 8>; static constexpr int rmin = 3; };
#endif
template <> struct PlanOf<256> { using type = Radices<16, 16>; static constexpr int rmin = 16; };
template <> struct PlanOf<512> { using type = Radices<8, 8, 8>; static constexpr int rmin = 8; };

struct Lines {
  int nl;   // number of lines
  int lw;   // lines per group
  int ls0;  // stride between lines inside a group
  int ls1;  // stride between groups
  int ps;   // stride between elements of a line
  __device__ __forceinline__ int base(int l) const { return (l / lw) * ls1 + (l % lw) * ls0; }
};

// tw[m] = exp(-2 pi i m / N) (forward convention); the inverse conjugates.
template <typename T2, int N, int R, bool FWD>
__device__ __forceinline__ void fft_pass(T2* lds, const Lines& L, int Ns, const T2* tw) {
  const int nb = L.nl * (N / R);
  const int b = threadIdx.x;
  const bool act = b < nb;
  T2 v[R];
  int base = 0, j = 0;
  if (act) {
    const int l = b % L.nl;
    j = b / L.nl;
    base = L.base(l);
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = lds[base + (j + r * (N / R)) * L.ps];
  }
  __syncthreads();
  if (act) {
    const int k = j % Ns;
    if (Ns > 1) {
      const int st = k * (N / (Ns * R));
#pragma unroll
      for (int r = 1; r < R; ++r) {
        T2 w = tw[r * st];
        if (!FWD) w.y = -w.y;
        v[r] = cmul(v[r], w);
      }
    }
    dftR<R, FWD>(v);
    const int d = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) lds[base + (d + r * Ns) * L.ps] = v[r];
  }
  __syncthreads();
}

template <typename T2, int N, bool FWD, int R0, int... Rs>
__device__ __forceinline__ void fft_chain(T2* lds, const Lines& L, const T2* tw, int Ns) {
  fft_pass<T2, N, R0, FWD>(lds, L, Ns, tw);
  if constexpr (sizeof...(Rs) > 0) fft_chain<T2, N, FWD, Rs...>(lds, L, tw, Ns * R0);
}

template <typename T2, int N, bool FWD, int... Rs>
__device__ __forceinline__ void fft_dispatch(T2* lds, const Lines& L, const T2* tw, Radices<Rs...>) {
  fft_chain<T2, N, FWD, Rs...>(lds, L, tw, 1);
}

// Transform all lines described by L; every thread of the block must call it.
template <typename T2, int N, bool FWD>
__device__ __forceinline__ void fft_lines(T2* lds, const Lines& L, const T2* tw) {
  fft_dispatch<T2, N, FWD>(lds, L, tw, typename PlanOf<N>::type{});
}

// Threads needed to transform nl lines of length N in one call.
template <int N> constexpr int fft_threads(int nl) { return nl * (N / PlanOf<N>::rmin); }

// Block size of the strided (y/x) passes: 16 lines per tile at most, 64..1024 threads.
// Used as the exact launch bound so small blocks get the full register file.
template <int N> constexpr int xy_threads() {
  return fft_threads<N>(16) < 64 ? 64 : (fft_threads<N>(16) > 1024 ? 1024 : fft_threads<N>(16));
}

// ----------------------------------------------------------------- global-edge variants
// Forward transforms use PlanOf<N> in order; inverse transforms use the reversed plan.
// The last forward pass of radix Rl leaves thread (l, j) holding positions
// j + r*N/Rl, which are exactly the operands of the first inverse pass (radix Rl,
// Ns = 1), so a pointwise k-space operation can sit in registers between them.
// Operands of the first pass are read straight from global memory through `ld`,
// results of the last pass are written straight to global memory through `st`;
// only the interior exchanges go through LDS.
template <typename A, typename B> struct RevImpl;
template <int... Done> struct RevImpl<Radices<>, Radices<Done...>> { using type = Radices<Done...>; };
template <int R0, int... Rs, int... Done> struct RevImpl<Radices<R0, Rs...>, Radices<Done...>> {
  using type = typename RevImpl<Radices<Rs...>, Radices<R0, Done...>>::type;
};
template <int N> using FwdPlan = typename PlanOf<N>::type;
template <int N> using InvPlan = typename RevImpl<typename PlanOf<N>::type, Radices<>>::type;

template <int R0, int... Rs> constexpr int first_radix(Radices<R0, Rs...>) { return R0; }
template <int R0> constexpr int last_radix(Radices<R0>) { return R0; }
template <int R0, int R1, int... Rs> constexpr int last_radix(Radices<R0, R1, Rs...>) {
  return last_radix(Radices<R1, Rs...>{});
}
// radix of the last forward pass (= first inverse pass)
template <int N> constexpr int edge_radix() { return last_radix(FwdPlan<N>{}); }

template <typename T2, int N, int R, bool FWD>
__device__ __forceinline__ void pass_math(T2* v, int j, int Ns, const T2* tw) {
  if (Ns > 1) {
    const int k = j % Ns;
    const int st = k * (N / (Ns * R));
#pragma unroll
    for (int r = 1; r < R; ++r) {
      T2 w = tw[r * st];
      if (!FWD) w.y = -w.y;
      v[r] = cmul(v[r], w);
    }
  }
  dftR<R, FWD>(v);
}

// One pass with selectable operand source and result sink.
// SRC: 0 = LDS, 1 = global via ld, 2 = registers vin.  DST: 0 = LDS, 1 = global via st, 2 = registers vout.
template <typename T2, int N, int R, bool FWD, int SRC, int DST, class LD, class ST>
__device__ __forceinline__ void gpass(T2* lds, const Lines& L, int Ns, const T2* tw, LD& ld, ST& st, T2* vreg,
                                      int b = threadIdx.x) {
  const int nb = L.nl * (N / R);
  const bool act = b < nb;
  const int l = act ? b % L.nl : 0;
  const int j = act ? b / L.nl : 0;
  const int base = L.base(l);
  T2 v[R];
  if (act) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int pos = j + r * (N / R);
      if constexpr (SRC == 0) v[r] = lds[base + pos * L.ps];
      else if constexpr (SRC == 1) v[r] = ld(l, pos);
      else v[r] = vreg[r];
    }
  }
  // LDS must not be overwritten while other threads still read it
  // LDS must not be overwritten while other threads still read it; an in-place
  // global -> global single pass must finish all reads before any write
  if constexpr (DST == 0 || (SRC == 1 && DST == 1)) __syncthreads();
  if (act) {
    pass_math<T2, N, R, FWD>(v, j, Ns, tw);
    const int d = (j / Ns) * Ns * R + j % Ns;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int pos = d + r * Ns;
      if constexpr (DST == 0) lds[base + pos * L.ps] = v[r];
      else if constexpr (DST == 1) st(l, pos, v[r]);
      else vreg[r] = v[r];
    }
  }
  if constexpr (DST == 0) __syncthreads();
}

template <typename T2, int N, bool FWD, int SRC, int DST, int Ns, class LD, class ST, int R0, int... Rs>
__device__ __forceinline__ void gchain(T2* lds, const Lines& L, const T2* tw, LD& ld, ST& st, T2* vreg,
                                       Radices<R0, Rs...>, int b = threadIdx.x) {
  if constexpr (sizeof...(Rs) == 0) {
    gpass<T2, N, R0, FWD, SRC, DST>(lds, L, Ns, tw, ld, st, vreg, b);
  } else {
    gpass<T2, N, R0, FWD, SRC, 0>(lds, L, Ns, tw, ld, st, vreg, b);
    gchain<T2, N, FWD, 0, DST, Ns * R0>(lds, L, tw, ld, st, vreg, Radices<Rs...>{}, b);
  }
}

// global -> global transform of the lines in L (LDS used for interior passes only)
template <typename T2, int N, bool FWD, class LD, class ST>
__device__ __forceinline__ void fft_g2g(T2* lds, const Lines& L, const T2* tw, LD ld, ST st) {
  T2* none = nullptr;
  if constexpr (FWD) gchain<T2, N, true, 1, 1, 1>(lds, L, tw, ld, st, none, FwdPlan<N>{});
  else gchain<T2, N, false, 1, 1, 1>(lds, L, tw, ld, st, none, InvPlan<N>{});
}

// global -> registers forward transform; vreg[edge_radix<N>()] receives positions j + r*N/Rl
template <typename T2, int N, class LD>
__device__ __forceinline__ void fft_g2r(T2* lds, const Lines& L, const T2* tw, LD ld, T2* vreg, int b = threadIdx.x) {
  auto st = [](int, int, T2) {};
  gchain<T2, N, true, 1, 2, 1>(lds, L, tw, ld, st, vreg, FwdPlan<N>{}, b);
}

// registers -> global inverse transform starting from the positions left by fft_g2r
template <typename T2, int N, class ST>
__device__ __forceinline__ void fft_r2g(T2* lds, const Lines& L, const T2* tw, T2* vreg, ST st, int b = threadIdx.x) {
  auto ld = [](int, int) { return T2{}; };
  // in-place use: every thread has read its operands (fft_g2r) before anyone writes
  __syncthreads();
  gchain<T2, N, false, 2, 1, 1>(lds, L, tw, ld, st, vreg, InvPlan<N>{}, b);
}

// The optional b of these calls is the calling thread's index (threadIdx.x by default). A caller
// whose work group is a power of two can pass threadIdx.x & (T - 1) so that the compiler knows every
// butterfly slot is taken (no branch around a pass), and an opaque copy of it per call so that the
// pass's index arithmetic is recomputed rather than kept live from one call to the next.
//
// The same pair with an explicit radix plan (forward in order, inverse reversed): vreg holds the
// last radix's positions j + r*N/Rl; the work group provides nl*N/rmin threads (passes of
// larger radix leave the surplus idle).
template <typename T2, int N, class Plan, class LD>
__device__ __forceinline__ void fft_g2r_p(T2* lds, const Lines& L, const T2* tw, LD ld, T2* vreg, int b = threadIdx.x) {
  auto st = [](int, int, T2) {};
  gchain<T2, N, true, 1, 2, 1>(lds, L, tw, ld, st, vreg, Plan{}, b);
}
template <typename T2, int N, class Plan, class ST>
__device__ __forceinline__ void fft_r2g_p(T2* lds, const Lines& L, const T2* tw, T2* vreg, ST st, int b = threadIdx.x) {
  auto ld = [](int, int) { return T2{}; };
  __syncthreads();
  gchain<T2, N, false, 2, 1, 1>(lds, L, tw, ld, st, vreg, typename RevImpl<Plan, Radices<>>::type{}, b);
}

// The same pair ending (forward) or starting (inverse) in the LDS tile instead of registers: the
// forward result sits at the positions the last pass's thread holds, (l, j + r*N/Rl), which the
// first inverse pass reads back, so a thread may work on its own positions in between without a
// barrier.
template <typename T2, int N, class Plan, class LD>
__device__ __forceinline__ void fft_g2l_p(T2* lds, const Lines& L, const T2* tw, LD ld, int b = threadIdx.x) {
  auto st = [](int, int, T2) {};
  T2* none = nullptr;
  gchain<T2, N, true, 1, 0, 1>(lds, L, tw, ld, st, none, Plan{}, b);
}
template <typename T2, int N, class Plan, class ST>
__device__ __forceinline__ void fft_l2g_p(T2* lds, const Lines& L, const T2* tw, ST st, int b = threadIdx.x) {
  auto ld = [](int, int) { return T2{}; };
  T2* none = nullptr;
  gchain<T2, N, false, 0, 1, 1>(lds, L, tw, ld, st, none, typename RevImpl<Plan, Radices<>>::type{}, b);
}

// All-LDS transform with the pass stride Ns known at compile time (constant index math);
// forward uses the plan in order, inverse the reversed plan.
template <typename T2, int N, bool FWD, int Ns, int R0, int... Rs>
__device__ __forceinline__ void fft_chain_c(T2* lds, const Lines& L, const T2* tw) {
  fft_pass<T2, N, R0, FWD>(lds, L, Ns, tw);
  if constexpr (sizeof...(Rs) > 0) fft_chain_c<T2, N, FWD, Ns * R0, Rs...>(lds, L, tw);
}
template <typename T2, int N, bool FWD, int... Rs>
__device__ __forceinline__ void fft_lines_c_(T2* lds, const Lines& L, const T2* tw, Radices<Rs...>) {
  fft_chain_c<T2, N, FWD, 1, Rs...>(lds, L, tw);
}
template <typename T2, int N, bool FWD>
__device__ __forceinline__ void fft_lines_c(T2* lds, const Lines& L, const T2* tw) {
  if constexpr (FWD) fft_lines_c_<T2, N, true>(lds, L, tw, FwdPlan<N>{});
  else fft_lines_c_<T2, N, false>(lds, L, tw, InvPlan<N>{});
}
// Same with an explicit radix plan (forward in order, inverse reversed).
template <typename T2, int N, bool FWD, class Plan>
__device__ __forceinline__ void fft_lines_p(T2* lds, const Lines& L, const T2* tw) {
  if constexpr (FWD) fft_lines_c_<T2, N, true>(lds, L, tw, Plan{});
  else fft_lines_c_<T2, N, false>(lds, L, tw, typename RevImpl<Plan, Radices<>>::type{});
}

// ------------------------------------------------------ passes for blocks of TB threads
// A pass of the all-LDS transform in a work group of TB threads whatever its butterfly count:
// thread b takes the butterflies b, b + TB, ... (at most K of them, K * TB >= L.nl * N / R).
template <typename T2, int N, int R, bool FWD, int Ns, int TB, int K>
__device__ __forceinline__ void fft_pass_tb(T2* lds, const Lines& L, const T2* tw) {
  const int nb = L.nl * (N / R);
  T2 v[K][R];
  int base[K], j[K];
#pragma unroll
  for (int q = 0; q < K; ++q) {
    const int b = (int)threadIdx.x + q * TB;
    j[q] = -1;
    base[q] = 0;
    if (b < nb) {
      j[q] = b / L.nl;
      base[q] = L.base(b % L.nl);
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = lds[base[q] + (j[q] + r * (N / R)) * L.ps];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < K; ++q) {
    if (j[q] < 0) continue;
    const int k = j[q] % Ns;
    if (Ns > 1) {
      const int st = k * (N / (Ns * R));
#pragma unroll
      for (int r = 1; r < R; ++r) {
        T2 w = tw[r * st];
        if (!FWD) w.y = -w.y;
        v[q][r] = cmul(v[q][r], w);
      }
    }
    dftR<R, FWD>(v[q]);
    const int d = (j[q] / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) lds[base[q] + (d + r * Ns) * L.ps] = v[q][r];
  }
  __syncthreads();
}
template <typename T2, int N, bool FWD, int Ns, int TB, int NL, int R0, int... Rs>
__device__ __forceinline__ void fft_chain_tb(T2* lds, const Lines& L, const T2* tw) {
  fft_pass_tb<T2, N, R0, FWD, Ns, TB, (NL * (N / R0) + TB - 1) / TB>(lds, L, tw);
  if constexpr (sizeof...(Rs) > 0) fft_chain_tb<T2, N, FWD, Ns * R0, TB, NL, Rs...>(lds, L, tw);
}
template <typename T2, int N, bool FWD, int TB, int NL, int... Rs>
__device__ __forceinline__ void fft_lines_tb_(T2* lds, const Lines& L, const T2* tw, Radices<Rs...>) {
  fft_chain_tb<T2, N, FWD, 1, TB, NL, Rs...>(lds, L, tw);
}
// ------------------------------------------------------ passes over one wave's lines
// Orders the calling wave's LDS accesses: its writes and reads before the call are complete and
// visible to all its lanes after it, and the compiler moves no LDS access across it.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// A pass of the all-LDS transform over lines that belong to the calling wave alone: lane l takes
// the butterflies l, l + 64, ... (at most K), numbered line-fastest (consecutive lanes on
// consecutive lines: lines that lie side by side) or, JF, butterfly-fastest (consecutive lanes
// along one line: lines whose elements lie side by side). No work-group barrier.
template <typename T2, int N, int R, bool FWD, int Ns, int K, bool JF, bool SK>
__device__ __forceinline__ void fft_wpass(T2* lds, const Lines& L, const T2* tw, int lane) {
  // SK: element p of a line sits at p + p / 16, and twiddle t at t + t / 16 (a pass that writes
  // or reads with a stride then spreads 16 lanes over 16 bank pairs)
  auto at = [&](int p) -> int { return (SK ? p + (p >> 4) : p) * L.ps; };
  constexpr int NB = N / R;  // butterflies per line
  const int nb = L.nl * NB;
  T2 v[K][R];
  int base[K], j[K];
#pragma unroll
  for (int q = 0; q < K; ++q) {
    const int b = lane + q * 64;
    j[q] = -1;
    base[q] = 0;
    if (K * 64 <= nb || b < nb) {
      j[q] = JF ? b % NB : b / L.nl;
      base[q] = L.base(JF ? b / NB : b % L.nl);
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = lds[base[q] + at(j[q] + r * (N / R))];
    }
  }
  wave_lds_sync();
#pragma unroll
  for (int q = 0; q < K; ++q) {
    if (j[q] < 0) continue;
    const int k = j[q] % Ns;
    if (Ns > 1) {
      const int st = k * (N / (Ns * R));
#pragma unroll
      for (int r = 1; r < R; ++r) {
        T2 w = tw[SK ? r * st + ((r * st) >> 4) : r * st];  // SK: the twiddle table is skewed too
        if (!FWD) w.y = -w.y;
        v[q][r] = cmul(v[q][r], w);
      }
    }
    dftR<R, FWD>(v[q]);
    const int d = (j[q] / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) lds[base[q] + at(d + r * Ns)] = v[q][r];
  }
  wave_lds_sync();
}
template <typename T2, int N, bool FWD, int Ns, int NL, bool JF, bool SK, int R0, int... Rs>
__device__ __forceinline__ void fft_wchain(T2* lds, const Lines& L, const T2* tw, int lane) {
  fft_wpass<T2, N, R0, FWD, Ns, (NL * (N / R0) + 63) / 64, JF, SK>(lds, L, tw, lane);
  if constexpr (sizeof...(Rs) > 0) fft_wchain<T2, N, FWD, Ns * R0, NL, JF, SK, Rs...>(lds, L, tw, lane);
}
template <typename T2, int N, bool FWD, int NL, bool JF, bool SK, int... Rs>
__device__ __forceinline__ void fft_wlines_(T2* lds, const Lines& L, const T2* tw, int lane, Radices<Rs...>) {
  fft_wchain<T2, N, FWD, 1, NL, JF, SK, Rs...>(lds, L, tw, lane);
}
// All lines of L (at most NL of them, touched by no other wave) with an explicit plan (forward in
// order, inverse reversed), by the calling wave; lane = its lane id; JF and SK as for fft_wpass.
template <typename T2, int N, bool FWD, class Plan, int NL, bool JF, bool SK = false>
__device__ __forceinline__ void fft_wlines(T2* lds, const Lines& L, const T2* tw, int lane) {
  if constexpr (FWD) fft_wlines_<T2, N, true, NL, JF, SK>(lds, L, tw, lane, Plan{});
  else fft_wlines_<T2, N, false, NL, JF, SK>(lds, L, tw, lane, typename RevImpl<Plan, Radices<>>::type{});
}

// All lines of L (at most NL of them) with an explicit plan (forward in order, inverse reversed)
// in a work group of TB threads; every thread of the group must call it.
template <typename T2, int N, bool FWD, class Plan, int TB, int NL>
__device__ __forceinline__ void fft_lines_tb(T2* lds, const Lines& L, const T2* tw) {
  if constexpr (FWD) fft_lines_tb_<T2, N, true, TB, NL>(lds, L, tw, Plan{});
  else fft_lines_tb_<T2, N, false, TB, NL>(lds, L, tw, typename RevImpl<Plan, Radices<>>::type{});
}

}  // namespace gsk
