// Contiguous-axis (z) passes of the FP32 fused step: real rows <-> half-spectrum rows.
//
// A tile is ZL consecutive rows (x*n + y) of every field. The complex M-point FFT
// (M = n/2) of the packed real row runs in LDS (row pitch M+1, compile-time tile
// shape, compile-time pass strides). The R2C / C2R split between the packed
// M-point spectrum and the n/2+1 half-spectrum values is fused with the global
// store / load: each thread owns the pair (k, M-k) of one row, reads or writes it
// straight from or to HBM, so the spectrum never takes an extra LDS round trip.
// Real rows are moved with 16-byte vector accesses (4 voxels per thread).
#pragma once
#include "kernels_common.h"

namespace gsk {

// rows per z tile and the exact block size (also used by the host)
__host__ __device__ constexpr int zrows(int M) { return M <= 32 ? 32 : (M <= 128 ? 16 : 8); }
__host__ __device__ constexpr int zthreads_of(int M, int rmin) {
  return zrows(M) * (M / rmin) < 64 ? 64 : (zrows(M) * (M / rmin) + 63) / 64 * 64;
}
template <int M> constexpr int zthreads() { return zthreads_of(M, PlanOf<M>::rmin); }

enum ZOp { Z_COPY = 0, Z_SQUARE = 1, Z_PREC_LOCAL = 2 };

// --------------------------------------------------------------------- split + global I/O
// half spectrum rows (global) -> packed M-point spectra Z (LDS), C2R direction
template <int M, int L>
__device__ __forceinline__ void z_load_c2r(float2* tile, const float2* __restrict__ kin, int F, const Geo& g,
                                           long r0, const float2* rt) {
  constexpr int H = M / 2 + 1;
  for (int e = threadIdx.x; e < F * L * H; e += blockDim.x) {
    const int k = e % H, fl = e / H, l = fl % L, f = fl / L;
    const float2* row = kin + kidx(g, f, r0 + l, 0);
    float2* z = tile + fl * (M + 1);
    if (k == 0) {
      const float a = row[0].x, b = row[M].x;
      z[0] = make_float2(a + b, a - b);
    } else {
      const float2 xk = row[k], xm = row[M - k];
      // E' = Xk + conj Xm ; O' = (Xk - conj Xm) * conj(W^k) ; Z = E' + i O'
      const float2 E = make_float2(xk.x + xm.x, xk.y - xm.y);
      const float2 D = make_float2(xk.x - xm.x, xk.y + xm.y);
      const float2 w = rt[k];
      const float2 O = cmul(D, make_float2(w.x, -w.y));
      z[k] = make_float2(E.x - O.y, E.y + O.x);
      if (k != M - k) z[M - k] = make_float2(E.x + O.y, -E.y + O.x);
    }
  }
  __syncthreads();
}

// packed M-point spectra Z (LDS) -> half spectrum rows (global), R2C direction
template <int M, int L>
__device__ __forceinline__ void z_store_r2c(const float2* tile, float2* __restrict__ kout, int F, const Geo& g,
                                            long r0, const float2* rt) {
  constexpr int H = M / 2 + 1;
  for (int e = threadIdx.x; e < F * L * H; e += blockDim.x) {
    const int k = e % H, fl = e / H, l = fl % L, f = fl / L;
    float2* row = kout + kidx(g, f, r0 + l, 0);
    const float2* z = tile + fl * (M + 1);
    if (k == 0) {
      const float2 z0 = z[0];
      row[0] = make_float2(z0.x + z0.y, 0.f);
      row[M] = make_float2(z0.x - z0.y, 0.f);
    } else {
      const float2 zk = z[k], zm = z[M - k];
      // E = (Zk + conj Zm)/2 ; O = (Zk - conj Zm)/(2i) ; X[k] = E + W^k O, X[M-k] = conj(E - W^k O)
      const float2 E = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
      const float2 O = make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
      const float2 WO = cmul(rt[k], O);
      row[k] = make_float2(E.x + WO.x, E.y + WO.y);
      if (k != M - k) row[M - k] = make_float2(E.x - WO.x, -(E.y - WO.y));
    }
  }
}

template <int M, int L, bool FWD>
__device__ __forceinline__ void z_fft(float2* tile, int F, const float2* tw) {
  const Lines Ls{L, L, M + 1, 0, 1};
  for (int f = 0; f < F; ++f) fft_lines_c<float2, M, FWD>(tile + f * L * (M + 1), Ls, tw);
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

// --------------------------------------------------------------------- forward z pass
// real fields -> pointwise op -> R2C -> half-spectrum rows.
// Martensite (Z_SQUARE): eta_p -> eta_p^2.  Copy (Z_COPY): unchanged.
// Precipitate (Z_PREC_LOCAL): fields (dc, eta_1..V) -> (mu_c, F_1..F_V), noise on eta.
template <int M, int OP>
__global__ void __launch_bounds__(zthreads<M>()) k_zfwd(const float* __restrict__ in, float2* __restrict__ kout, Geo g,
                                                       int F, MartK mk, PrecK pk, const float2* __restrict__ twM,
                                                       const float2* __restrict__ rtn) {
  constexpr int L = zrows(M), Q = M / 2;  // Q quads of 4 voxels per row
  extern __shared__ float2 smem[];
  float2* tw = smem;
  float2* rt = smem + M;
  float2* tile = smem + 2 * M;
  const long r0 = (long)blockIdx.x * L;
  const int n = 2 * M;
  z_twiddles<M>(tw, rt, twM, rtn);
  if constexpr (OP == Z_PREC_LOCAL) {
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
      for (int i = 0; i < 4; ++i) {
        const float c = pk.cbar + v[0][i];
        float s2 = 0.f;
        for (int a = 1; a <= V; ++a) s2 += v[a][i] * v[a][i];
        out[0][i] = 0.5f * pk.A2 * s2 + pk.A1 * (c - pk.c1);
      }
      for (int a = 1; a <= V; ++a) {
        float xi[4] = {0.f, 0.f, 0.f, 0.f};
        if (pk.noise != 0.f) normal4(pk.seed, (uint32_t)(vox >> 2), (uint32_t)(a - 1), pk.step, xi);
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
        }
      }
      for (int f = 0; f < F; ++f) {
        float2* row = tile + (f * L + l) * (M + 1);
        row[2 * q] = make_float2(out[f][0], out[f][1]);
        row[2 * q + 1] = make_float2(out[f][2], out[f][3]);
      }
    }
  } else {
    for (int e = threadIdx.x; e < F * L * Q; e += blockDim.x) {
      const int q = e % Q, fl = e / Q, l = fl % L, f = fl / L;
      float4 t = *reinterpret_cast<const float4*>(in + f * g.nvox + (r0 + l) * n + 4 * q);
      if (OP == Z_SQUARE) t = make_float4(t.x * t.x, t.y * t.y, t.z * t.z, t.w * t.w);
      float2* row = tile + fl * (M + 1);
      row[2 * q] = make_float2(t.x, t.y);
      row[2 * q + 1] = make_float2(t.z, t.w);
    }
  }
  __syncthreads();
  z_fft<M, L, true>(tile, F, tw);
  z_store_r2c<M, L>(tile, kout, F, g, r0, rt);
  (void)mk;
}

// --------------------------------------------------------------------- martensite middle z pass
// R_hat rows -> C2R -> F = f'(eta) + 2 w eta R + a xi (Philox in registers) -> R2C.
// VT > 0: the variant count is a compile-time constant and eta is prefetched into
// registers before the spectra are loaded; VT = 0: any variant count, eta read late.
template <int M, int VT>
__global__ void __launch_bounds__(zthreads<M>()) k_zmid_mart(float2* __restrict__ kbuf, const float* __restrict__ eta,
                                                            Geo g, MartK mk, const float2* __restrict__ twM,
                                                            const float2* __restrict__ rtn) {
  constexpr int L = zrows(M), Q = M / 2, ZT = zthreads<M>();
  constexpr int QPT = (L * Q + ZT - 1) / ZT;  // quads per thread
  constexpr int VR = VT > 0 ? VT : 1;
  extern __shared__ float2 smem[];
  float2* tw = smem;
  float2* rt = smem + M;
  float2* tile = smem + 2 * M;
  const long r0 = (long)blockIdx.x * L;
  const int F = VT > 0 ? VT : mk.F, n = 2 * M;
  float4 pre[QPT][VR];
  if constexpr (VT > 0) {
#pragma unroll
    for (int i = 0; i < QPT; ++i) {
      const int e = threadIdx.x + i * ZT;
      if (e < L * Q) {
        const long vox = (r0 + e / Q) * n + 4 * (e % Q);
#pragma unroll
        for (int f = 0; f < VR; ++f) pre[i][f] = *reinterpret_cast<const float4*>(eta + f * g.nvox + vox);
      }
    }
  }
  z_twiddles<M>(tw, rt, twM, rtn);
  z_load_c2r<M, L>(tile, kbuf, F, g, r0, rt);
  z_fft<M, L, false>(tile, F, tw);
#pragma unroll
  for (int i = 0; i < QPT; ++i) {
    const int e = threadIdx.x + i * ZT;
    if (e >= L * Q) continue;
    const int l = e / Q, q = e % Q;
    const long vox = (r0 + l) * n + 4 * q;
    float et[8][4];
    for (int f = 0; f < F; ++f) {
      const float4 t = VT > 0 ? pre[i][f < VR ? f : 0] : *reinterpret_cast<const float4*>(eta + f * g.nvox + vox);
      et[f][0] = t.x; et[f][1] = t.y; et[f][2] = t.z; et[f][3] = t.w;
    }
    float s2[4];
    for (int k = 0; k < 4; ++k) {
      float s = 0.f;
      for (int f = 0; f < F; ++f) s += et[f][k] * et[f][k];
      s2[k] = s;
    }
    for (int f = 0; f < F; ++f) {
      float2* row = tile + (f * L + l) * (M + 1);
      const float2 ra = row[2 * q], rb = row[2 * q + 1];
      const float R[4] = {ra.x, ra.y, rb.x, rb.y};
      float xi[4] = {0.f, 0.f, 0.f, 0.f};
      if (mk.noise != 0.f) normal4(mk.seed, (uint32_t)(vox >> 2), (uint32_t)f, mk.step, xi);
      float o[4];
      for (int k = 0; k < 4; ++k) {
        const float e1 = et[f][k];
        o[k] = mk.A * e1 - mk.B * e1 * e1 + mk.C * e1 * s2[k] + mk.two_wel * e1 * R[k] + mk.noise * xi[k];
      }
      row[2 * q] = make_float2(o[0], o[1]);
      row[2 * q + 1] = make_float2(o[2], o[3]);
    }
  }
  __syncthreads();
  z_fft<M, L, true>(tile, F, tw);
  z_store_r2c<M, L>(tile, kbuf, F, g, r0, rt);
}

// --------------------------------------------------------------------- inverse z pass
// half-spectrum rows -> C2R -> real fields.
template <int M>
__global__ void __launch_bounds__(zthreads<M>()) k_zinv(const float2* __restrict__ kin, float* __restrict__ out, Geo g,
                                                       int F, const float2* __restrict__ twM,
                                                       const float2* __restrict__ rtn) {
  constexpr int L = zrows(M), Q = M / 2;
  extern __shared__ float2 smem[];
  float2* tw = smem;
  float2* rt = smem + M;
  float2* tile = smem + 2 * M;
  const long r0 = (long)blockIdx.x * L;
  const int n = 2 * M;
  z_twiddles<M>(tw, rt, twM, rtn);
  z_load_c2r<M, L>(tile, kin, F, g, r0, rt);
  z_fft<M, L, false>(tile, F, tw);
  for (int e = threadIdx.x; e < F * L * Q; e += blockDim.x) {
    const int q = e % Q, fl = e / Q, l = fl % L, f = fl / L;
    const float2* row = tile + fl * (M + 1);
    const float2 a = row[2 * q], b = row[2 * q + 1];
    *reinterpret_cast<float4*>(out + f * g.nvox + (r0 + l) * n + 4 * q) = make_float4(a.x, a.y, b.x, b.y);
  }
}

}  // namespace gsk
