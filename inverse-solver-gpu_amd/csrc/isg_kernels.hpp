// isg_kernels.hpp -- HIP kernels of the MLFMA operator and the BiCGSTAB
// vector algebra, written for gfx950 (64-lane wavefronts, 256-thread blocks).
//
// Operator stages (reference counterparts in parentheses):
//   k_p2m     leaf radiation patterns            (kernels.cu:6-30)
//   k_m2m     child -> parent interpolate+shift  (kernels.cu:83-136)
//   k_m2l     diagonal translation               (kernels.cu:239-266)
//   k_l2l     parent -> child shift+anterpolate  (kernels.cu:137-238)
//   k_l2p     leaf receiving patterns            (kernels.cu:32-81)
//   k_p2p     near field                         (kernels.cu:292-319)
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

// r[b*64+n] = sum_k basisT[k][n] loc[b][k]   (assigns)
template <typename T, typename TO>
__global__ void __launch_bounds__(kBlock) k_l2p(const cplx<T>* __restrict__ agl, long as, cplx<TO>* __restrict__ r,
                                                long rs, const cplx<T>* __restrict__ basisT, int S, int B,
                                                const int* __restrict__ act) {
  int rhs = blockIdx.y;
  if (rhs_off(act, rhs)) return;
  long idx = (long)blockIdx.x * kBlock + threadIdx.x;
  if (idx >= (long)B * 64) return;
  int b = (int)(idx >> 6), n = (int)(idx & 63);
  const cplx<T>* lb = agl + rhs * as + (long)b * S;
  cplx<T> acc = mkc<T>(0, 0);
  for (int k = 0; k < S; k++) acc = cfma(basisT[k * 64 + n], lb[k], acc);
  r[rhs * rs + idx] = ccast<TO>(acc);
}

// r[b*64+m] += sum over in-domain 3x3 neighbour boxes nb (slot s) of
//              sum_n near[s][n][m] y[nb*64+n]
template <typename T, typename TI>
__global__ void __launch_bounds__(kBlock) k_p2p(const cplx<TI>* __restrict__ y, long ys, cplx<TI>* __restrict__ r,
                                                long rs, const cplx<T>* __restrict__ near, int nbs, int B,
                                                const int* __restrict__ act) {
  int rhs = blockIdx.y;
  if (rhs_off(act, rhs)) return;
  long idx = (long)blockIdx.x * kBlock + threadIdx.x;
  if (idx >= (long)B * 64) return;
  int b = (int)(idx >> 6), m = (int)(idx & 63);
  int X = (int)compact2((unsigned)b), Y = (int)compact2((unsigned)b >> 1);
  cplx<T> acc = mkc<T>(0, 0);
  for (int s = 0; s < 9; s++) {
    int sx = X + s % 3 - 1, sy = Y + s / 3 - 1;
    if (sx < 0 || sy < 0 || sx >= nbs || sy >= nbs) continue;
    long nb = morton2((unsigned)sx, (unsigned)sy);
    const cplx<TI>* yb = y + rhs * ys + nb * 64;
    const cplx<T>* blk = near + s * 4096 + m;
#pragma unroll 8
    for (int n = 0; n < 64; n++) acc = cfma(blk[n * 64], ccast<T>(yb[n]), acc);
  }
  cplx<TI>* dst = r + rhs * rs + idx;
  *dst = cadd(*dst, ccast<TI>(acc));
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
