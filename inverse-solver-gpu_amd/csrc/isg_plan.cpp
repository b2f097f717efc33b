// isg_plan.cpp -- quadtree geometry and MLFMA operator tables (host, FP64).
// See isg_plan.hpp for the layouts.  Reference mathematics: setup_mlfma.cpp
// and integrate.cpp (cited per function below).
#include "isg_plan.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>

namespace isg {

namespace {

// Primitive of log(x^2 + y^2) used by the self-term integral (integrate.cpp:11-13).
double log_prim(double x, double y) {
  return -3 * x * y + x * y * std::log(x * x + y * y) + x * x * std::atan(y / x) +
         y * y * std::atan(x / y);
}

// Lagrange weights for resampling a periodic pattern from `nsrc` to `ndst`
// equispaced angles: target sample m uses the `taps` source samples centred at
// ceil(m * nsrc / ndst) (setup_mlfma.cpp:573-599 and :652-677).  The
// interpolation variable is the un-wrapped normalised angle; indices wrap.
// Returns the first (wrapped) source index.
int lagrange_weights(int m, int ndst, int nsrc, int taps, double* w) {
  double ratio = (double)nsrc / ndst;
  int centre = (taps % 2 == 0) ? (int)std::ceil(m * ratio) : (int)std::round(m * ratio);
  double xm = (double)m / ndst;
  for (int n = 0; n < taps; n++) {
    double xn = (double)(centre + n - taps / 2) / nsrc;
    double prod = 1;
    for (int k = 0; k < taps; k++) {
      if (k == n) continue;
      double xk = (double)(centre + k - taps / 2) / nsrc;
      prod = prod * (xm - xk) / (xn - xk);
    }
    w[n] = prod;
  }
  int first = (centre - taps / 2) % nsrc;
  if (first < 0) first += nsrc;
  return first;
}

// Transposed interpolator: for every destination-grid sample j of the source
// level, the list of (target sample, weight * scale) pairs in ascending
// target order, weight scaled by num/den (setup_mlfma.cpp:687-693).
void transpose_lists(int ndst, int nsrc, int taps, const std::vector<double>& wT, const std::vector<int>& first,
                     int num, int den, std::vector<std::vector<std::pair<int, double>>>& lists) {
  lists.assign(nsrc, {});
  for (int m = 0; m < ndst; m++)
    for (int t = 0; t < taps; t++) {
      int j = (first[m] + t) % nsrc;
      lists[j].push_back({m, wT[(size_t)t * ndst + m] * num / den});
    }
}

}  // namespace

cd green_pixel(double k0, double res, cd pt, cd pb) {
  double d = std::abs(pt - pb);
  cd val(j0(k0 * d), 0.0), anal(0.0, 0.0);
  if (d < res / 8) {
    // small-argument model of Y0 integrated over the pixel (integrate.cpp:23-35)
    val = val + cd(0.0, 0.5772156649015329 * 2 / M_PI + 2 / M_PI * std::log(0.5));
    cd dd = pb - pt;
    double h = res / 2;
    double I = log_prim(dd.real() + h, dd.imag() + h) - log_prim(dd.real() - h, dd.imag() + h) -
               log_prim(dd.real() + h, dd.imag() - h) + log_prim(dd.real() - h, dd.imag() - h);
    anal = cd(0.0, (I / 2 + std::log(k0) * res * res) * 2 / M_PI);
  } else {
    val = val + cd(0.0, y0(k0 * d));  // midpoint rule (integrate.cpp:37-38)
  }
  val = val * res;
  val = val * res;
  return (val + anal) * cd(0.0, 0.25);
}

bool build_plan(const PlanConfig& cfg, Plan& P) {
  P = Plan();
  P.cfg = cfg;
  if (cfg.box != 8) { P.error = "box must be 8 (64 pixels per leaf box)"; return false; }
  if (cfg.level < 3 || cfg.level > 12) { P.error = "level must be in [3, 12]"; return false; }
  if (cfg.ninter < 2 || cfg.ninter % 2 || cfg.ninter > 64) { P.error = "ninter must be even in [2, 64]"; return false; }
  if (cfg.numrx < 1) { P.error = "numrx must be >= 1"; return false; }
  if (!(cfg.res > 0) || !(cfg.scale > 0)) { P.error = "res and scale must be positive"; return false; }
  if (cfg.beta < 1 || cfg.beta > 32) { P.error = "beta must be in [1, 32]"; return false; }

  const int L = cfg.level, ni = cfg.ninter;
  P.L = L;
  P.k0 = 2 * M_PI * cfg.scale;
  P.side = cfg.box << (L - 1);
  P.N = P.side * P.side;
  P.B = 1 << (2 * (L - 1));
  P.nclus.resize(L);
  P.nsamp.resize(L);
  P.nterm.resize(L);
  P.csize.resize(L);
  // sample counts: excess-bandwidth rule (setup_mlfma.cpp:75-85)
  for (int i = 0; i < L; i++) {
    P.nclus[i] = 1 << (2 * i);
    P.csize[i] = cfg.res * cfg.box * std::pow(2.0, L - i - 1);
    double kr = std::sqrt(2.0) * P.k0 * P.csize[i];
    double t = kr + 1.8 * std::pow((double)cfg.beta, 2.0 / 3) * std::pow(kr, 1.0 / 3);
    P.nterm[i] = (int)std::floor(t);
    P.nsamp[i] = 4 * P.nterm[i];
  }
  const double D = P.csize[0], res = cfg.res, k0 = P.k0;
  const int Sl = P.nsamp[L - 1];

  // pixel centres: leaf box (X, Y) in Morton order, pixel (k, l) row-major in
  // the box, domain centred at the origin (setup_mlfma.cpp:155-185)
  const double a_leaf = P.csize[L - 1];
  P.pos.resize(P.N);
  P.raster2tree.resize(P.N);
  const int bs = 1 << (L - 1);
  for (int b = 0; b < P.B; b++) {
    int X = morton_compact((uint32_t)b), Y = morton_compact((uint32_t)b >> 1);
    for (int k = 0; k < 8; k++)
      for (int l = 0; l < 8; l++) {
        int n = b * 64 + k * 8 + l;
        P.pos[n] = cd(-D / 2 + X * a_leaf + (l + 0.5) * res, D / 2 - Y * a_leaf - (k + 0.5) * res);
        P.raster2tree[(Y * 8 + k) * P.side + X * 8 + l] = n;
      }
  }
  (void)bs;

  // near-field blocks: slot s = 3 (dY + 1) + (dX + 1) for the source box at
  // (X + dX, Y + dY); entry [basis n][test m] (setup_mlfma.cpp:459-505)
  const bool ht = cfg.host_tables;
  if (ht) P.near.assign((size_t)9 * 64 * 64, cd(0, 0));
  for (int s = 0; s < 9 && ht; s++) {
    int dY = s / 3 - 1, dX = s % 3 - 1;
    for (int n = 0; n < 64; n++)
      for (int m = 0; m < 64; m++) {
        int kt = m / 8, lt = m % 8, ks = n / 8, ls = n % 8;
        cd pt((lt + 0.5) * res, -(kt + 0.5) * res);
        cd pb((dX * 8 + ls + 0.5) * res, -(dY * 8 + ks + 0.5) * res);
        P.near[((size_t)s * 64 + n) * 64 + m] = green_pixel(k0, res, pt, pb);
      }
  }
  {
    int nb = 1 << (L - 1);
    P.near_pairs = 0;
    for (int Y = 0; Y < nb; Y++)
      for (int X = 0; X < nb; X++)
        for (int dY = -1; dY <= 1; dY++)
          for (int dX = -1; dX <= 1; dX++)
            if (X + dX >= 0 && X + dX < nb && Y + dY >= 0 && Y + dY < nb) P.near_pairs++;
  }

  // P2M radiation and L2P receiving patterns about the box centre
  // (integrate.cpp:43-69, setup_mlfma.cpp:546-553)
  if (ht) {
    P.coeffT.assign((size_t)64 * Sl, cd(0, 0));
    P.basisT.assign((size_t)Sl * 64, cd(0, 0));
  }
  for (int n = 0; n < 64 && ht; n++) {
    cd d((n % 8 + 0.5) * res - a_leaf / 2, a_leaf / 2 - (n / 8 + 0.5) * res);
    for (int k = 0; k < Sl; k++) {
      double ang = 2 * M_PI * k / Sl;
      double ud = std::cos(ang) * d.real() + std::sin(ang) * d.imag();
      cd c = std::exp(cd(0.0, -k0 * ud));
      c = c * res;
      P.coeffT[(size_t)n * Sl + k] = c * res;
      P.basisT[(size_t)k * 64 + n] = std::exp(cd(0.0, k0 * ud)) / cd(Sl, 0) * cd(0.0, 0.25);
    }
  }

  // shifts between a parent centre and its children (setup_mlfma.cpp:556-570)
  P.shiftM.assign(L, {});
  P.shiftL.assign(L, {});
  for (int i = 0; i < L - 1 && ht; i++) {
    int S = P.nsamp[i];
    double q = P.csize[i] / 4;
    P.shiftM[i].resize(4 * S);
    P.shiftL[i].resize(4 * S);
    for (int c = 0; c < 4; c++) {
      int k_ = c / 2, l_ = c % 2;
      cd rho(q * (1 - 2 * l_), -q * (1 - 2 * k_));  // parent centre minus child centre
      for (int k = 0; k < S; k++) {
        double ang = 2 * M_PI * k / S;
        double pr = k0 * std::cos(ang) * rho.real() + k0 * std::sin(ang) * rho.imag();
        P.shiftM[i][c * S + k] = std::exp(cd(0.0, pr));
        P.shiftL[i][c * S + k] = std::exp(cd(0.0, -pr));
      }
    }
  }

  // interpolators child (i+1) -> parent (i) and anterpolators parent -> child
  // (setup_mlfma.cpp:573-631)
  P.interpT.assign(L, {});
  P.interp0.assign(L, {});
  P.anterpT.assign(L, {});
  P.anterp0.assign(L, {});
  std::vector<double> w(ni);
  for (int i = 0; i < L - 1; i++) {
    int Sp = P.nsamp[i], Sc = P.nsamp[i + 1];
    P.interpT[i].assign((size_t)ni * Sp, 0.0);
    P.interp0[i].assign(Sp, 0);
    for (int m = 0; m < Sp; m++) {
      P.interp0[i][m] = lagrange_weights(m, Sp, Sc, ni, w.data());
      for (int t = 0; t < ni; t++) P.interpT[i][(size_t)t * Sp + m] = w[t];
    }
  }
  const int na = 2 * ni;
  for (int i = 1; i < L; i++) {
    int Sc = P.nsamp[i], Sp = P.nsamp[i - 1];
    std::vector<std::vector<std::pair<int, double>>> lists;
    // anterpolation weight = interpolation weight * S_child / S_parent
    lists.assign(Sc, {});
    for (int m = 0; m < Sp; m++)
      for (int t = 0; t < ni; t++) {
        int j = (P.interp0[i - 1][m] + t) % Sc;
        lists[j].push_back({m, P.interpT[i - 1][(size_t)t * Sp + m] * Sc / Sp});
      }
    P.anterpT[i].assign((size_t)na * Sc, 0.0);
    P.anterp0[i].assign(Sc, 0);
    for (int j = 0; j < Sc; j++) {
      auto& lj = lists[j];
      if ((int)lj.size() > na) { P.error = "anterpolator needs more than 2*ninter taps"; return false; }
      // rotate to start after the wrap-around so indices run contiguously
      for (size_t n = 1; n < lj.size(); n++)
        if (lj[n].first != lj[n - 1].first + 1) {
          std::rotate(lj.begin(), lj.begin() + n, lj.end());
          break;
        }
      int first = lj.empty() ? 0 : lj[0].first;
      for (size_t n = 0; n < lj.size(); n++) {
        if (lj[n].first != (first + (int)n) % Sp) { P.error = "anterpolator taps not contiguous"; return false; }
        P.anterpT[i][n * Sc + j] = lj[n].second;
      }
      P.anterp0[i][j] = first;
    }
  }

  // diagonal translation operators on the 7x7 offset grid, entry 7(dY+3)+(dX+3)
  // for a source box offset (dX, dY) from the target (setup_mlfma.cpp:506-543)
  P.trans.assign(L, {});
  P.far_pairs.assign(L, 0);
  std::vector<cd> ipow(P.nterm[0] + 2);
  for (int p = 0; p <= P.nterm[0]; p++) ipow[p] = std::exp(cd(0.0, (p + 1) * M_PI / 2));
  for (int i = 2; i < L; i++) {
    int S = P.nsamp[i], Pn = P.nterm[i];
    double a = P.csize[i];
    if (ht) P.trans[i].assign((size_t)49 * S, cd(0, 0));
    std::vector<cd> h(Pn + 1);
    for (int dY = -3; dY <= 3 && ht; dY++)
      for (int dX = -3; dX <= 3; dX++) {
        if (std::abs(dX) <= 1 && std::abs(dY) <= 1) continue;
        cd d(-dX * a, dY * a);  // target centre minus source centre
        double r = std::abs(d), phi = std::arg(d);
        for (int p = 0; p <= Pn; p++) h[p] = cd(jn(p, k0 * r), yn(p, k0 * r));
        cd* T = &P.trans[i][(size_t)(7 * (dY + 3) + (dX + 3)) * S];
        for (int k = 0; k < S; k++) {
          cd acc = h[0];
          for (int p = 1; p <= Pn; p++) acc = acc + 2 * std::cos(p * (2 * M_PI * k / S - phi)) * h[p] * ipow[p - 1];
          T[k] = acc;
        }
      }
    int nb = 1 << i;
    long long cnt = 0;
    for (int Y = 0; Y < nb; Y++)
      for (int X = 0; X < nb; X++)
        for (int sY = (Y / 2 - 1) * 2; sY < (Y / 2 + 2) * 2; sY++)
          for (int sX = (X / 2 - 1) * 2; sX < (X / 2 + 2) * 2; sX++) {
            if (sX < 0 || sY < 0 || sX >= nb || sY >= nb) continue;
            if (std::abs(sX - X) <= 1 && std::abs(sY - Y) <= 1) continue;
            cnt++;
          }
    P.far_pairs[i] = cnt;
  }

  // receivers: interpolation of the root pattern to numrx equispaced angles
  // and its scaled transpose (setup_mlfma.cpp:642-695)
  int R = cfg.numrx, S0 = P.nsamp[0];
  P.rxT.assign((size_t)ni * R, 0.0);
  P.rx0.assign(R, 0);
  for (int m = 0; m < R; m++) {
    P.rx0[m] = lagrange_weights(m, R, S0, ni, w.data());
    for (int t = 0; t < ni; t++) P.rxT[(size_t)t * R + m] = w[t];
  }
  {
    std::vector<std::vector<std::pair<int, double>>> lists;
    transpose_lists(R, S0, ni, P.rxT, P.rx0, S0, R, lists);
    P.rxaT.assign((size_t)na * S0, 0.0);
    P.rxaI.assign((size_t)na * S0, 0);
    for (int j = 0; j < S0; j++) {
      if ((int)lists[j].size() > na) { P.error = "receiver anterpolator needs more than 2*ninter taps (numrx too large)"; return false; }
      for (size_t n = 0; n < lists[j].size(); n++) {
        P.rxaI[n * S0 + j] = lists[j][n].first;
        P.rxaT[n * S0 + j] = lists[j][n].second;
      }
    }
  }
  return true;
}

}  // namespace isg
