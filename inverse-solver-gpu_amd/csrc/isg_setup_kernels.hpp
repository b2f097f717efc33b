// isg_setup_kernels.hpp -- operator set-up on the device (FP64).
//
// The tables of setup_mlfma.cpp:453-631 that dominate its cost, computed by
// kernels straight into the device layouts of isg_plan.hpp:
//   near blocks     9 x 64 x 64 pixel-integrated Green's function values
//                   (integrate.cpp:11-41, setup_mlfma.cpp:459-505)
//   P2M / L2P       radiation and receiving patterns of the leaf pixels
//                   (integrate.cpp:43-69, setup_mlfma.cpp:546-553)
//   shifts          parent <-> child plane-wave phase factors (setup_mlfma.cpp:556-570)
//   translations    T(k) = H0 + sum_p 2 cos(p (phi_k - arg d)) i^p Hp(k0 |d|) on the
//                   7 x 7 offset grid of every level >= 2 (setup_mlfma.cpp:506-543)
// The Hankel functions Hp = Jp + i Yp of one offset come from one thread: Y by
// forward recurrence from Y0, Y1 (stable for increasing order), J by Miller's
// backward recurrence normalised with J0 + 2 sum_k J2k = 1.
#pragma once

#include <hip/hip_runtime.h>

#include "isg_device.hpp"

namespace isg {

namespace setup {

constexpr int kSB = 256;  // threads per block

// primitive of log(x^2 + y^2) over a rectangle corner (integrate.cpp:11-13)
__device__ __forceinline__ double log_prim(double x, double y) {
  return -3 * x * y + x * y * log(x * x + y * y) + x * x * atan(y / x) + y * y * atan(x / y);
}

// (i/4) H0^(1) integrated over the source pixel at pb seen from pt: midpoint
// rule, or the small-argument model of Y0 for the self term (integrate.cpp:15-41)
__device__ zd green_pixel(double k0, double res, double ptx, double pty, double pbx, double pby) {
  const double dx = pbx - ptx, dy = pby - pty;
  const double d = sqrt(dx * dx + dy * dy);
  double vr = j0(k0 * d), vi, ai = 0;
  if (d < res / 8) {
    vi = 0.5772156649015329 * 2 / M_PI + 2 / M_PI * log(0.5);
    const double h = res / 2;
    const double I = log_prim(dx + h, dy + h) - log_prim(dx - h, dy + h) - log_prim(dx + h, dy - h) +
                     log_prim(dx - h, dy - h);
    ai = (I / 2 + log(k0) * res * res) * 2 / M_PI;
  } else {
    vi = y0(k0 * d);
  }
  vr = vr * res * res;
  vi = vi * res * res + ai;
  return mkc<double>(-0.25 * vi, 0.25 * vr);  // (vr + i vi) * i / 4
}

// near[s][n][m], s = 3 (dY + 1) + (dX + 1): test pixel m of the target box,
// basis pixel n of the source box at (X + dX, Y + dY)
__global__ void __launch_bounds__(kSB) k_near_table(double k0, double res, zd* __restrict__ near) {
  const int idx = blockIdx.x * kSB + threadIdx.x;
  if (idx >= 9 * 64 * 64) return;
  const int m = idx & 63, n = (idx >> 6) & 63, s = idx >> 12;
  const int dY = s / 3 - 1, dX = s % 3 - 1;
  const int kt = m >> 3, lt = m & 7, ks = n >> 3, ls = n & 7;
  near[idx] = green_pixel(k0, res, (lt + 0.5) * res, -(kt + 0.5) * res, (dX * 8 + ls + 0.5) * res,
                          -(dY * 8 + ks + 0.5) * res);
}

// coeffT[n][k] = res^2 exp(-i k0 u_k . d_n), basisT[k][n] = (i/4) / S exp(+i k0 u_k . d_n),
// d_n the pixel's offset from its leaf-box centre
__global__ void __launch_bounds__(kSB) k_p2m_table(double k0, double res, double a_leaf, int S,
                                                   zd* __restrict__ coeffT, zd* __restrict__ basisT) {
  const int idx = blockIdx.x * kSB + threadIdx.x;
  if (idx >= 64 * S) return;
  const int k = idx % S, n = idx / S;
  const double ddx = (n % 8 + 0.5) * res - a_leaf / 2, ddy = a_leaf / 2 - (n / 8 + 0.5) * res;
  double sa, ca;
  sincos(2 * M_PI * k / S, &sa, &ca);
  const double ph = k0 * (ca * ddx + sa * ddy);
  double sp, cp;
  sincos(ph, &sp, &cp);
  coeffT[(long)n * S + k] = mkc<double>(cp * res * res, -sp * res * res);
  basisT[(long)k * 64 + n] = mkc<double>(-0.25 * sp / S, 0.25 * cp / S);
}

// shiftM[c][k] = exp(+i k0 u_k . rho_c), shiftL = its conjugate; rho_c = parent
// centre minus child c's centre, children c = 2 k + l (row k from the top)
__global__ void __launch_bounds__(kSB) k_shift_table(double k0, double q, int S, zd* __restrict__ sm,
                                                     zd* __restrict__ sl) {
  const int idx = blockIdx.x * kSB + threadIdx.x;
  if (idx >= 4 * S) return;
  const int k = idx % S, c = idx / S;
  const double rx = q * (1 - 2 * (c % 2)), ry = -q * (1 - 2 * (c / 2));
  double sa, ca;
  sincos(2 * M_PI * k / S, &sa, &ca);
  double sp, cp;
  sincos(k0 * (ca * rx + sa * ry), &sp, &cp);
  sm[idx] = mkc<double>(cp, sp);
  sl[idx] = mkc<double>(cp, -sp);
}

// H_p(x), p = 0..P, of offset o (7 x 7 grid, entry 7 (dY + 3) + (dX + 3)) of a
// level with box size a: h[o][p].  One thread per offset; the 9 near offsets
// are skipped.  jwork holds M + 2 doubles per offset for the backward
// recurrence (M > max(P, x) the start order).
__global__ void __launch_bounds__(64) k_hankel(double k0, double a, int P, int M, double* __restrict__ jwork,
                                              zd* __restrict__ h) {
  const int o = blockIdx.x * 64 + threadIdx.x;
  if (o >= 49) return;
  const int dY = o / 7 - 3, dX = o % 7 - 3;
  if (abs(dX) <= 1 && abs(dY) <= 1) return;
  const double x = k0 * a * sqrt((double)(dX * dX + dY * dY));
  zd* ho = h + (long)o * (P + 1);
  // Y: forward recurrence Y_{p+1} = (2p / x) Y_p - Y_{p-1}
  double ym = y0(x), yc = y1(x);
  ho[0].y = ym;
  if (P >= 1) ho[1].y = yc;
  for (int p = 1; p < P; p++) {
    const double yn = (2.0 * p / x) * yc - ym;
    ym = yc;
    yc = yn;
    ho[p + 1].y = yc;
  }
  // J: backward recurrence J_{p-1} = (2p / x) J_p - J_{p+1} from J_{M+1} = 0, J_M = 1,
  // rescaled when it grows, normalised by J0 + 2 sum J_{2k} = 1
  double* jw = jwork + (long)o * (M + 2);
  double jp1 = 0.0, jc = 1.0, sum = 0.0;
  jw[M] = jc;
  for (int p = M; p >= 1; p--) {
    const double jm = (2.0 * p / x) * jc - jp1;
    jp1 = jc;
    jc = jm;
    jw[p - 1] = jc;
    if (p - 1 > 0 && ((p - 1) & 1) == 0) sum += 2 * jc;
    if (fabs(jc) > 1e250) {  // rescale everything computed so far
      for (int q = p - 1; q <= M; q++) jw[q] *= 1e-250;
      jp1 *= 1e-250;
      jc *= 1e-250;
      sum *= 1e-250;
    }
  }
  sum += jw[0];
  const double inv = 1.0 / sum;
  for (int p = 0; p <= P; p++) ho[p].x = jw[p] * inv;
}

// T[o][k] = H_0 + sum_{p=1}^{P} 2 cos(p (2 pi k / S - phi_o)) i^p H_p, phi_o = arg(d_o),
// d_o = (-dX a, dY a) the target centre minus the source centre
__global__ void __launch_bounds__(kSB) k_trans_table(int S, int P, double a, const zd* __restrict__ h,
                                                     zd* __restrict__ T) {
  const int k = blockIdx.x * kSB + threadIdx.x;
  const int o = blockIdx.y;
  if (k >= S) return;
  const int dY = o / 7 - 3, dX = o % 7 - 3;
  zd* out = T + (long)o * S + k;
  if (abs(dX) <= 1 && abs(dY) <= 1) {
    *out = mkc<double>(0.0, 0.0);
    return;
  }
  const double phi = atan2(dY * a, -dX * a);
  const double th = 2 * M_PI * k / S - phi;
  const zd* ho = h + (long)o * (P + 1);
  double ar = ho[0].x, ai = ho[0].y;
  for (int p = 1; p <= P; p++) {
    const double c = 2 * cos(p * th);
    const zd hp = ho[p];
    // i^p hp: rotate by p quarter turns
    double qr, qi;
    switch (p & 3) {
      case 0: qr = hp.x; qi = hp.y; break;
      case 1: qr = -hp.y; qi = hp.x; break;
      case 2: qr = -hp.x; qi = -hp.y; break;
      default: qr = hp.y; qi = -hp.x; break;
    }
    ar += c * qr;
    ai += c * qi;
  }
  *out = mkc<double>(ar, ai);
}

// FP32 copy of a complex FP64 table
__global__ void __launch_bounds__(kSB) k_table_f32(const zd* __restrict__ in, cplx<float>* __restrict__ out,
                                                   long n) {
  for (long i = (long)blockIdx.x * kSB + threadIdx.x; i < n; i += (long)gridDim.x * kSB)
    out[i] = mkc<float>((float)in[i].x, (float)in[i].y);
}

}  // namespace setup

}  // namespace isg
