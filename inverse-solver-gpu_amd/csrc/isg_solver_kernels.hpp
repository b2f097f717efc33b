// isg_solver_kernels.hpp -- BiCGSTAB vector algebra for gfx950.
//
// The reference runs every BiCGSTAB step as a separate saxpy / prep / post
// kernel and finishes each dot product on the host after a blocking copy
// (bicgs_gpu.cu:16-58, 116-224; kernels.cu:334-413).  Here each step is one
// fused pass that also produces the dot products the next scalar needs:
//   k_bi_init   r = b - A x, r~ = r, partial <b,b>, <r,r>
//   k_bi_pupd   p = r + beta (p - omega v), y = prep(p)
//   (near)      v = post(p, G y), partial <r~,v>    -- epilogue of k_near_mfma
//   k_bi_supd   s = r - alpha v, y = prep(s)
//   (near)      t = post(s, G y), partial <t,s>, <t,t>    -- likewise
//   k_bi_xrupd  x += alpha p; x += omega s; r = s - omega t; partial <r,r>, <r~,r>
// Partial sums are FP64, one per block, reduced in a fixed order by
// k_bi_fin, which also advances the per-right-hand-side scalars on the
// device (no host round trip).  prep/post are the contrast products of
// kernels.cu:394-409: forward y = g x, v = x - Gy; adjoint y = conj x,
// v = x - conj(g Gy).
#pragma once

#include "isg_device.hpp"
#include "isg_engine.hpp"
#include "isg_kernels.hpp"

namespace isg {

enum { PH_INIT = 0, PH_ALPHA = 1, PH_OMEGA = 2, PH_RESID = 3 };
// r0 (may be null) supplies a residual computed beforehand (blended start).
template <typename V>
__global__ void __launch_bounds__(kBlock) k_bi_init(const zd* __restrict__ x, const cplx<V>* __restrict__ gy,
                                                    const zd* __restrict__ g, long gs, const zd* __restrict__ b,
                                                    const cplx<V>* __restrict__ r0, cplx<V>* __restrict__ r,
                                                    cplx<V>* __restrict__ rt, long N, int herm, int xzero,
                                                    double2* __restrict__ part) {
  int rhs = blockIdx.y;
  long o = rhs * N;
  zd acc[2] = {mkc<double>(0, 0), mkc<double>(0, 0)};
  ISG_VEC_LOOP(N) {
    zd bv = b[o + i];
    zd rv;
    if (r0) {
      rv = ldv(r0, o + i);
    } else {
      zd ax = xzero ? mkc<double>(0, 0) : post_val(x[o + i], ldv(gy, o + i), g, rhs * gs + i, herm);
      rv = csub(bv, ax);
    }
    if (r) rv = stv(r, o + i, rv);  // r, rt null: the norms only (residual check)
    if (rt) stv(rt, o + i, rv);
    acc[0].x += bv.x * bv.x + bv.y * bv.y;
    acc[1].x += rv.x * rv.x + rv.y * rv.y;
  }
  block_store<2>(acc, part + (long)rhs * kSlots * gridDim.x, gridDim.x, blockIdx.x);
}

template <typename V>
__global__ void __launch_bounds__(kBlock) k_bi_pupd(const cplx<V>* __restrict__ r, cplx<V>* __restrict__ p,
                                                    const cplx<V>* __restrict__ v, const zd* __restrict__ g, long gs,
                                                    cplx<V>* __restrict__ y, long N, int herm,
                                                    const BiState* __restrict__ st) {
  int rhs = blockIdx.y;
  const BiState& S = st[rhs];
  if (!S.active) return;
  long o = rhs * N;
  zd beta = S.beta, mom = mkc<double>(-S.omega.x, -S.omega.y);
  int first = S.first;
  ISG_VEC_LOOP(N) {
    zd pv;
    if (first) {
      pv = ldv(r, o + i);
    } else {
      zd tmp = cfma(mom, ldv(v, o + i), ldv(p, o + i));
      pv = cfma(beta, tmp, ldv(r, o + i));
    }
    pv = stv(p, o + i, pv);
    stv(y, o + i, prep_val(pv, g, rhs * gs + i, herm));
  }
}

template <typename V>
__global__ void __launch_bounds__(kBlock) k_bi_supd(const cplx<V>* __restrict__ r, const cplx<V>* __restrict__ v,
                                                    const zd* __restrict__ g, long gs, cplx<V>* __restrict__ s,
                                                    cplx<V>* __restrict__ y, long N, int herm,
                                                    const BiState* __restrict__ st) {
  int rhs = blockIdx.y;
  const BiState& S = st[rhs];
  if (!S.active) return;
  long o = rhs * N;
  zd mal = mkc<double>(-S.alpha.x, -S.alpha.y);
  ISG_VEC_LOOP(N) {
    zd sv = stv(s, o + i, cfma(mal, ldv(v, o + i), ldv(r, o + i)));
    stv(y, o + i, prep_val(sv, g, rhs * gs + i, herm));
  }
}

template <typename V>
__global__ void __launch_bounds__(kBlock) k_bi_xrupd(zd* __restrict__ x, const cplx<V>* __restrict__ p,
                                                     const cplx<V>* __restrict__ s, const cplx<V>* __restrict__ t,
                                                     cplx<V>* __restrict__ r, const cplx<V>* __restrict__ rt, long N,
                                                     const BiState* __restrict__ st, double2* __restrict__ part) {
  int rhs = blockIdx.y;
  const BiState& S = st[rhs];
  if (!S.active) return;
  long o = rhs * N;
  zd al = S.alpha, om = S.omega, mom = mkc<double>(-S.omega.x, -S.omega.y);
  zd acc[2] = {mkc<double>(0, 0), mkc<double>(0, 0)};
  ISG_VEC_LOOP(N) {
    zd sv = ldv(s, o + i);
    zd xv = cfma(al, ldv(p, o + i), x[o + i]);
    xv = cfma(om, sv, xv);
    x[o + i] = xv;
    zd rv = stv(r, o + i, cfma(mom, ldv(t, o + i), sv));
    zd a = ldv(rt, o + i);
    acc[0].x += rv.x * rv.x + rv.y * rv.y;
    acc[1].x += a.x * rv.x + a.y * rv.y;
    acc[1].y += a.x * rv.y - a.y * rv.x;
  }
  block_store<2>(acc, part + (long)rhs * kSlots * gridDim.x, gridDim.x, blockIdx.x);
}

// One block per right-hand side: fixed-order reduction of the partials, then
// the scalar step of bicgs_gpu.cu:148-213 for that right-hand side.
__global__ void __launch_bounds__(kBlock) k_bi_fin(int phase, BiState* __restrict__ st, int* __restrict__ active,
                                                   const double2* __restrict__ part, int nblk,
                                                   const double* __restrict__ tolv, int maxit) {
  int rhs = blockIdx.x;
  BiState& S = st[rhs];
  if (phase != PH_INIT && !S.active) return;
  const double tol = (phase == PH_INIT) ? tolv[rhs] : S.tol;
  const double2* pr = part + (long)rhs * kSlots * nblk;
  int ns = (phase == PH_ALPHA) ? 1 : 2;
  __shared__ double sh[4][kBlock / 64];
  double v[4] = {0, 0, 0, 0};
  for (int s = 0; s < ns; s++)
    for (int j = threadIdx.x; j < nblk; j += kBlock) {
      double2 q = pr[s * nblk + j];
      v[2 * s] += q.x;
      v[2 * s + 1] += q.y;
    }
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int s = 0; s < 4; s++) {
    double a = wave_sum(v[s]);
    if (lane == 0) sh[s][w] = a;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double tot[4] = {0, 0, 0, 0};
  for (int s = 0; s < 4; s++)
    for (int i = 0; i < kBlock / 64; i++) tot[s] += sh[s][i];
  zd s0 = mkc<double>(tot[0], tot[1]), s1 = mkc<double>(tot[2], tot[3]);
  if (phase == PH_INIT) {
    double bn = sqrt(s0.x);
    if (bn < 1e-20) bn = 1;
    S.tol = tol;
    S.bnrm = bn;
    S.err = sqrt(s1.x) / bn;
    S.iter = 0;
    S.first = 1;
    S.beta = mkc<double>(0, 0);
    S.alpha = mkc<double>(0, 0);
    S.omega = mkc<double>(0, 0);
    if (S.err > tol) {
      S.rho = mkc<double>(s1.x, 0.0);
      S.iter = 1;
      if (zabs(S.rho) < 1e-200) {
        S.status = 2;
        S.active = 0;
      } else {
        S.status = 1;
        S.active = 1;
      }
    } else {
      S.status = (S.err < tol) ? 0 : 1;
      S.active = 0;
    }
  } else if (phase == PH_ALPHA) {
    S.alpha = zdiv(S.rho, s0);
  } else if (phase == PH_OMEGA) {
    double tt = s1.x;
    S.omega = mkc<double>(s0.x / tt, s0.y / tt);
    if (zabs(s0) < 1e-200 && tt < 1e-200) S.omega = mkc<double>(1.0, 0.0);
  } else {
    S.err = sqrt(s0.x) / S.bnrm;
    if (S.err < tol) {
      S.status = 0;
      S.active = 0;
    } else if (zabs(S.omega) < 1e-200) {
      S.status = 3;
      S.active = 0;
    } else if (S.iter >= maxit) {
      S.status = 1;
      S.active = 0;
    } else {
      S.rho1 = S.rho;
      S.rho = s1;
      S.iter++;
      S.first = 0;
      if (zabs(S.rho) < 1e-200) {
        S.status = 2;
        S.active = 0;
      } else {
        S.beta = cmul(zdiv(S.rho, S.rho1), zdiv(S.alpha, S.omega));
      }
    }
  }
  active[rhs] = S.active;
}

// Blended start: ad = A d (post of the operator product), partials
// <ad, r> (= sum conj(ad) r), <ad, ad> and <r, r>.
template <typename V>
__global__ void __launch_bounds__(kBlock) k_bl_dots(const zd* __restrict__ d, const cplx<V>* __restrict__ gy,
                                                    const zd* __restrict__ g, long gs, const cplx<V>* __restrict__ r,
                                                    cplx<V>* __restrict__ ad, long N, int herm,
                                                    double2* __restrict__ part) {
  int rhs = blockIdx.y;
  long o = rhs * N;
  zd acc[2] = {mkc<double>(0, 0), mkc<double>(0, 0)};
  ISG_VEC_LOOP(N) {
    zd a = stv(ad, o + i, post_val(d[o + i], ldv(gy, o + i), g, rhs * gs + i, herm));
    zd rv = ldv(r, o + i);
    acc[0].x += a.x * rv.x + a.y * rv.y;
    acc[0].y += a.x * rv.y - a.y * rv.x;
    acc[1].x += a.x * a.x + a.y * a.y;
    acc[1].y += rv.x * rv.x + rv.y * rv.y;
  }
  block_store<2>(acc, part + (long)rhs * kSlots * gridDim.x, gridDim.x, blockIdx.x);
}

// lambda = <ad, r> / <ad, ad> per right-hand side, fixed-order sums; 0 when
// ad = 0 or the sums are not finite (a correction d holding NaN or Inf is
// then ignored: the solve starts from x alone).  ratio (may be null) gets
// |r - lambda ad| / |r|, the blend's reduction of the initial residual.
__global__ void __launch_bounds__(kBlock) k_bl_fin(const double2* __restrict__ part, int nblk, zd* __restrict__ lam,
                                                   double* __restrict__ ratio) {
  int rhs = blockIdx.x;
  const double2* pr = part + (long)rhs * kSlots * nblk;
  __shared__ double sh[4][kBlock / 64];
  double v[4] = {0, 0, 0, 0};
  for (int j = threadIdx.x; j < nblk; j += kBlock) {
    v[0] += pr[j].x;
    v[1] += pr[j].y;
    v[2] += pr[nblk + j].x;
    v[3] += pr[nblk + j].y;
  }
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int s = 0; s < 4; s++) {
    double a = wave_sum(v[s]);
    if (lane == 0) sh[s][w] = a;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double t[4] = {0, 0, 0, 0};
  for (int s = 0; s < 4; s++)
    for (int i = 0; i < kBlock / 64; i++) t[s] += sh[s][i];
  zd l = mkc<double>(0, 0);
  if (t[2] > 0 && isfinite(t[2])) l = mkc<double>(t[0] / t[2], t[1] / t[2]);
  if (!(isfinite(l.x) && isfinite(l.y))) l = mkc<double>(0, 0);
  lam[rhs] = l;
  if (ratio) {
    // |r - lambda ad|^2 = |r|^2 - |<ad, r>|^2 / <ad, ad> at the minimising lambda
    double q = 1.0;
    if ((l.x != 0 || l.y != 0) && t[3] > 0) q = sqrt(fmax(0.0, 1.0 - (t[0] * t[0] + t[1] * t[1]) / (t[2] * t[3])));
    ratio[rhs] = q;
  }
}

// x += lambda d, r -= lambda ad
template <typename V>
__global__ void __launch_bounds__(kBlock) k_bl_apply(zd* __restrict__ x, const zd* __restrict__ d,
                                                     cplx<V>* __restrict__ r, const cplx<V>* __restrict__ ad,
                                                     const zd* __restrict__ lam, long N) {
  int rhs = blockIdx.y;
  long o = rhs * N;
  zd l = lam[rhs], ml = mkc<double>(-l.x, -l.y);
  if (l.x == 0 && l.y == 0) return;  // nothing to add (and d may hold non-finite values)
  ISG_VEC_LOOP(N) {
    x[o + i] = cfma(l, d[o + i], x[o + i]);
    stv(r, o + i, cfma(ml, ldv(ad, o + i), ldv(r, o + i)));
  }
}

// x += c  (defect correction of the mixed-precision solve)
__global__ void __launch_bounds__(kBlock) k_add(zd* __restrict__ x, const zd* __restrict__ c, long N) {
  int rhs = blockIdx.y;
  long o = rhs * N;
  ISG_VEC_LOOP(N) x[o + i] = cadd(x[o + i], c[o + i]);
}

// y = prep(x)
template <typename V>
__global__ void __launch_bounds__(kBlock) k_prep(const zd* __restrict__ x, const zd* __restrict__ g, long gs,
                                                 cplx<V>* __restrict__ y, long N, int herm) {
  int rhs = blockIdx.y;
  long o = rhs * N;
  ISG_VEC_LOOP(N) stv(y, o + i, prep_val(x[o + i], g, rhs * gs + i, herm));
}

}  // namespace isg
