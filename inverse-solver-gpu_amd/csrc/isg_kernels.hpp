// isg_kernels.hpp -- HIP kernels of the MLFMA operator and the BiCGSTAB
// vector algebra, written for gfx950 (64-lane wavefronts, 256-thread blocks).
//
// Operator stages (reference counterparts in parentheses):
//   k_p2m     leaf radiation patterns            (kernels.cu:6-30)
//   k_m2m     child -> parent interpolate+shift  (kernels.cu:83-136)
//   k_m2l     diagonal translation               (kernels.cu:239-266)
//   k_l2l     parent -> child shift+anterpolate  (kernels.cu:137-238)
//   k_near_mfma  near field + leaf receiving patterns on the matrix cores
//                                                (kernels.cu:32-81, :292-319)
//   k_rx_*    receiver interpolation / anterpolation (matvec_gpu.cu:351-361, :395-401)
// Grids are (work items / 256, right-hand sides); blockIdx.y selects the
// right-hand side and an optional `act` mask skips converged ones.
#pragma once

#include "isg_device.hpp"

namespace isg {

constexpr int kBlock = 256;

__device__ __forceinline__ bool rhs_off(const int* act, int rhs) { return act != nullptr && act[rhs] == 0; }

// agm[b][k] = sum_n coeffT[n][k] y[b*64+n]
template <typename T, typename TI>
__global__ void __launch_bounds__(kBlock) k_p2m(const cplx<TI>* __restrict__ y, long ys, cplx<T>* __restrict__ agm,
                                                long as, const cplx<T>* __restrict__ coeffT, int S, int B,
                                                const int* __restrict__ act) {
  int rhs = blockIdx.y;
  if (rhs_off(act, rhs)) return;
  long idx = (long)blockIdx.x * kBlock + threadIdx.x;
  if (idx >= (long)B * S) return;
  int b = (int)(idx / S), k = (int)(idx - (long)b * S);
  const cplx<TI>* yb = y + rhs * ys + (long)b * 64;
  cplx<T> acc = mkc<T>(0, 0);
#pragma unroll 8
  for (int n = 0; n < 64; n++) acc = cfma(coeffT[n * S + k], ccast<T>(yb[n]), acc);
  agm[rhs * as + idx] = acc;
}

// parent[p][k] = sum_c shift[c][k] * sum_t w[t][k] child[4p+c][(i0[k]+t) mod Sc]
template <typename T>
__global__ void __launch_bounds__(kBlock) k_m2m(const cplx<T>* __restrict__ child, long cs, cplx<T>* __restrict__ par,
                                                long ps, const T* __restrict__ wT, const int* __restrict__ i0,
                                                const cplx<T>* __restrict__ shift, int Sp, int Sc, int np, int ntap,
                                                const int* __restrict__ act) {
  int rhs = blockIdx.y;
  if (rhs_off(act, rhs)) return;
  long idx = (long)blockIdx.x * kBlock + threadIdx.x;
  if (idx >= (long)np * Sp) return;
  int p = (int)(idx / Sp), k = (int)(idx - (long)p * Sp);
  int j0 = i0[k];
  cplx<T> acc = mkc<T>(0, 0);
  for (int c = 0; c < 4; c++) {
    const cplx<T>* ch = child + rhs * cs + (long)(4 * p + c) * Sc;
    cplx<T> red = mkc<T>(0, 0);
    int j = j0;
    for (int t = 0; t < ntap; t++) {
      T w = wT[t * Sp + k];
      cplx<T> v = ch[j];
      red.x += w * v.x;
      red.y += w * v.y;
      j = (j + 1 == Sc) ? 0 : j + 1;
    }
    acc = cfma(red, shift[c * Sp + k], acc);
  }
  par[rhs * ps + idx] = acc;
}

// loc[m][k] = sum over the interaction list of m of T[offset][k] agm[n][k].
// The list is enumerated from Morton coordinates: children of the parent's
// 3x3 neighbours that are not themselves neighbours of m.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_m2l(const cplx<T>* __restrict__ agm, cplx<T>* __restrict__ agl, long st,
                                                const cplx<T>* __restrict__ trans, int lev, int S,
                                                const int* __restrict__ act) {
  int rhs = blockIdx.y;
  if (rhs_off(act, rhs)) return;
  int nb = 1 << lev;
  long tot = (long)nb * nb * S;
  long idx = (long)blockIdx.x * kBlock + threadIdx.x;
  if (idx >= tot) return;
  int m = (int)(idx / S), k = (int)(idx - (long)m * S);
  int X = (int)compact2((unsigned)m), Y = (int)compact2((unsigned)m >> 1);
  int x0 = (X >> 1) * 2 - 2, y0 = (Y >> 1) * 2 - 2;
  const cplx<T>* src = agm + rhs * st;
  cplx<T> acc = mkc<T>(0, 0);
  for (int sy = y0; sy < y0 + 6; sy++) {
    if (sy < 0 || sy >= nb) continue;
    int dy = sy - Y;
    for (int sx = x0; sx < x0 + 6; sx++) {
      if (sx < 0 || sx >= nb) continue;
      int dx = sx - X;
      if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1) continue;
      long n = morton2((unsigned)sx, (unsigned)sy);
      acc = cfma(trans[(7 * (dy + 3) + dx + 3) * S + k], src[n * S + k], acc);
    }
  }
  agl[rhs * st + idx] = acc;
}

// child[c][k] (+)= sum_t a[t][k] * shiftL[c%4][j] * par[c/4][j], j = (j0[k]+t) mod Sp
template <typename T>
__global__ void __launch_bounds__(kBlock) k_l2l(const cplx<T>* __restrict__ par, long ps, cplx<T>* __restrict__ child,
                                                long cs, const T* __restrict__ aT, const int* __restrict__ j0s,
                                                const cplx<T>* __restrict__ shift, int Sp, int Sc, int np, int ntap,
                                                int accumulate, const int* __restrict__ act) {
  int rhs = blockIdx.y;
  if (rhs_off(act, rhs)) return;
  long idx = (long)blockIdx.x * kBlock + threadIdx.x;
  if (idx >= (long)4 * np * Sc) return;
  int c = (int)(idx / Sc), k = (int)(idx - (long)c * Sc);
  int p = c >> 2, cm = c & 3;
  const cplx<T>* pv = par + rhs * ps + (long)p * Sp;
  const cplx<T>* sh = shift + cm * Sp;
  cplx<T> acc = mkc<T>(0, 0);
  int j = j0s[k];
  for (int t = 0; t < ntap; t++) {
    T w = aT[t * Sc + k];
    cplx<T> v = cmul(sh[j], pv[j]);
    acc.x += w * v.x;
    acc.y += w * v.y;
    j = (j + 1 == Sp) ? 0 : j + 1;
  }
  cplx<T>* dst = child + rhs * cs + idx;
  *dst = accumulate ? cadd(*dst, acc) : acc;
}

// ---------------------------------------------------------------------------
// Matrix-core (MFMA) near field.  For a tile of NB Morton-consecutive leaf
// boxes the output r[box][m] (64 rows m = test pixel, NB columns) is one
// complex GEMM over K = 9*64 (near blocks times the gathered neighbour
// pixels) + S (leaf receiving pattern times the leaf locals, i.e. L2P), so
// the near field and L2P are fused and r is written once.
// Wave w owns rows 16w..16w+15; each 16x16 output block accumulates four real
// products (re*re, -im*im, re*im, im*re) with v_mfma_*_16x16x4.  Operand
// layout of the 16x16x4 instructions (checked on gfx950): A[i][k] in lane
// i + 16k, B[k][j] in lane j + 16k; the 16x16 result D[i][j] is in lane
// j + 16(i/4), register i%4 for f32 and in lane j + 16(i%4), register i/4
// for f64 -- Mfma<T>::row gives the row of (lane group g, register q).
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mac(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int q) { return g + 4 * q; }
};
template <>
struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mac(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int q) { return 4 * g + q; }
};

// LDS row length (complex elements) of the near/L2P tile: 64 pixels or S
// locals, plus one element of padding so 16 lanes reading 16 different rows
// hit 16 different banks.
__host__ __device__ inline int near_row(int S) { return (S > 64 ? S : 64) + 1; }

template <typename T, typename TI, int NB>
__global__ void __launch_bounds__(kBlock) k_near_mfma(const cplx<TI>* __restrict__ y, long ys,
                                                      const cplx<T>* __restrict__ loc, long ls,
                                                      cplx<TI>* __restrict__ r, long rs,
                                                      const cplx<T>* __restrict__ near,
                                                      const cplx<T>* __restrict__ basisT, int S, int nbs,
                                                      int with_near, const int* __restrict__ act) {
  typedef typename Mfma<T>::v4 v4;
  constexpr int NN = NB / 16;
  extern __shared__ char smem_raw[];
  cplx<T>* tile = reinterpret_cast<cplx<T>*>(smem_raw);
  const int row = near_row(S);
  int rhs = blockIdx.y;
  if (rhs_off(act, rhs)) return;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const long b0 = (long)blockIdx.x * NB;
  v4 cr[NN], ci[NN];
#pragma unroll
  for (int j = 0; j < NN; j++) {
    cr[j] = v4{0, 0, 0, 0};
    ci[j] = v4{0, 0, 0, 0};
  }
  const cplx<TI>* yr = y + rhs * ys;
  if (with_near) {
    for (int s = 0; s < 9; s++) {
      const int dX = s % 3 - 1, dY = s / 3 - 1;
      __syncthreads();
      for (int idx = tid; idx < NB * 64; idx += kBlock) {
        int j = idx >> 6, n = idx & 63;
        unsigned b = (unsigned)(b0 + j);
        int sx = (int)compact2(b) + dX, sy = (int)compact2(b >> 1) + dY;
        cplx<T> v = mkc<T>(0, 0);
        if (sx >= 0 && sy >= 0 && sx < nbs && sy < nbs) v = ccast<T>(yr[(long)morton2(sx, sy) * 64 + n]);
        tile[j * row + n] = v;
      }
      __syncthreads();
      const cplx<T>* A = near + (long)s * 4096 + 16 * w + li;
#pragma unroll 4
      for (int kk = 0; kk < 16; kk++) {
        const int k = 4 * kk + lk;
        cplx<T> a = A[k * 64];
        T nai = -a.y;
#pragma unroll
        for (int j = 0; j < NN; j++) {
          cplx<T> bv = tile[(16 * j + li) * row + k];
          cr[j] = Mfma<T>::mac(a.x, bv.x, cr[j]);
          cr[j] = Mfma<T>::mac(nai, bv.y, cr[j]);
          ci[j] = Mfma<T>::mac(a.x, bv.y, ci[j]);
          ci[j] = Mfma<T>::mac(a.y, bv.x, ci[j]);
        }
      }
    }
  }
  // leaf locals -> pixels (L2P) as the last K block
  __syncthreads();
  const cplx<T>* lr = loc + rhs * ls + b0 * S;
  for (int idx = tid; idx < NB * S; idx += kBlock) {
    int j = idx / S, k = idx - j * S;
    tile[j * row + k] = lr[idx];
  }
  __syncthreads();
  const cplx<T>* A = basisT + 16 * w + li;
  for (int kk = 0; kk < S / 4; kk++) {
    const int k = 4 * kk + lk;
    cplx<T> a = A[k * 64];
    T nai = -a.y;
#pragma unroll
    for (int j = 0; j < NN; j++) {
      cplx<T> bv = tile[(16 * j + li) * row + k];
      cr[j] = Mfma<T>::mac(a.x, bv.x, cr[j]);
      cr[j] = Mfma<T>::mac(nai, bv.y, cr[j]);
      ci[j] = Mfma<T>::mac(a.x, bv.y, ci[j]);
      ci[j] = Mfma<T>::mac(a.y, bv.x, ci[j]);
    }
  }
  cplx<TI>* out = r + rhs * rs;
#pragma unroll
  for (int j = 0; j < NN; j++) {
    long box = b0 + 16 * j + li;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int m = 16 * w + Mfma<T>::row(lk, q);
      out[box * 64 + m] = mkc<TI>((TI)cr[j][q], (TI)ci[j][q]);
    }
  }
}

// receiver values: out[m] = sum_t wT[t][m] agm0[(i0[m]+t) mod S0]
template <typename T>
__global__ void __launch_bounds__(kBlock) k_rx_interp(const cplx<T>* __restrict__ agm0, long as, zd* __restrict__ out,
                                                      long os, const T* __restrict__ wT, const int* __restrict__ i0,
                                                      int S0, int R, int ntap) {
  int rhs = blockIdx.y;
  int m = blockIdx.x * kBlock + threadIdx.x;
  if (m >= R) return;
  const cplx<T>* a = agm0 + rhs * as;
  cplx<T> acc = mkc<T>(0, 0);
  int j = i0[m];
  for (int t = 0; t < ntap; t++) {
    T w = wT[t * R + m];
    acc.x += w * a[j].x;
    acc.y += w * a[j].y;
    j = (j + 1 == S0) ? 0 : j + 1;
  }
  out[rhs * os + m] = ccast<double>(acc);
}

// root locals from receiver values: agl0[k] = (sum_t aT[t][k] in[aI[t][k]]) * (-4i) * sc
template <typename T>
__global__ void __launch_bounds__(kBlock) k_rx_anterp(const zd* __restrict__ in, long is, cplx<T>* __restrict__ agl0,
                                                      long as, const T* __restrict__ aT, const int* __restrict__ aI,
                                                      int S0, int ntap, double sc) {
  int rhs = blockIdx.y;
  int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= S0) return;
  const zd* v = in + rhs * is;
  zd acc = mkc<double>(0, 0);
  for (int t = 0; t < ntap; t++) {
    double w = (double)aT[t * S0 + k];
    zd u = v[aI[t * S0 + k]];
    acc.x += w * u.x;
    acc.y += w * u.y;
  }
  zd r = cmul(acc, mkc<double>(0.0, -4.0));
  r = mkc<double>(r.x * sc, r.y * sc);
  agl0[rhs * as + k] = ccast<T>(r);
}

}  // namespace isg
