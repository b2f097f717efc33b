// isg_oracle.cpp -- CPU restatement of the reference MLFMA / BiCGSTAB / DBIM path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it (through
// oracle/oracle.py).  The product library (inverse-solver-gpu_amd/, libisg.so)
// never links, loads or calls anything here, and must not.
//
// It restates the algorithm of merthidayetoglu/inverse-solver-GPU's CPU path
// in plain FP64 C++ with OpenMP.  Array layouts follow the reference's host
// arrays (setup_mlfma.cpp), so per-stage checkpoints (multipoles / locals per
// level) compare element by element with the GPU library after a layout map.
// Every output element is reduced by one thread in a fixed order, so results
// do not depend on the OpenMP thread count.  Built -O3 without fast-math.
//
// Semantics taken where the reference's two solver variants disagree: the
// BiCGSTAB follows the GPU variant that the reference's Makefile actually
// builds (bicgs_gpu.cu:116-224): breakdown thresholds 1e-200 and the omega
// NaN guard of bicgs_gpu.cu:195-197.  The CPU variant (bicgs.cpp:88-179)
// differs only in the 1e-20 thresholds and has no guard; it produces NaN on
// the first DBIM derivative solve (contrast zero), which the guard prevents.
//
// How each function is pinned (tests/test_oracle_*.py):
//   green_pixel        scipy hankel1 (off-diagonal), 2-D quadrature of the
//                      reference's small-argument model (self term)
//   hn                 scipy jv/yv evaluated in the test (tests/test_oracle_green.py)
//   tree / lists       brute-force geometry (distances, Morton order, raster map)
//   near blocks        direct evaluation of green_pixel on the box pair
//   trans              addition theorem: plane-wave sum vs H0(k|r-r'|)
//   interp / anterp    Lagrange identities (partition of unity, polynomial
//                      exactness), transpose relation, contiguity of taps
//   coeff / basis      closed-form exponentials
//   mlfma              dense direct MoM product (relative error 1e-4..1e-6)
//   mlfmah             conj(G conj x) identity against mlfma
//   aggregate          far-field closed form: sum_n res^2 exp(-i k u.r_n) x_n
//   aggregateh         adjoint identity <aggregate x, y> vs <x, aggregateh y>
//                      up to the reference's scale -4i numrx res^2 (closed form)
//   bicgs              numpy.linalg.solve on the dense operator
//   dbim               gradient (with the per-transmitter Tikhonov term) by
//                      central finite differences of the regularised misfit;
//                      step length alpha against a finite-difference Frechet
//                      derivative (minimiser of the linearised misfit);
//                      Polak-Ribiere beta by conjugacy of successive
//                      directions in the Born limit (tests/test_oracle_dbim.py).
//                      Unpinned: where the conjugate sits in beta's numerator
//                      (the imaginary part a quadratic misfit cannot see)
//   phantoms           closed forms (blob peak, ellipse membership)
// No function is left without such a check.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>
#include <omp.h>

typedef std::complex<double> cd;

namespace {

struct Oracle {
  // Configuration: mom.cpp:4-32 (res, box, level, k0, numrx) and
  // setup_mlfma.cpp:3-4 (beta, ninter).
  int L = 0, box = 8, beta = 5, ninter = 30, numrx = 0;
  double res = 0.1, k0 = 0.0;
  // 0: plane-wave transmitters and far-field receivers (the reference);
  // > 0: line-source transmitters and point receivers on a circle of this
  // radius (wavelengths), equispaced from angle 0
  double radius = 0.0;
  int N = 0, len = 0;
  std::vector<int> nclus, nsamp, nterm;
  std::vector<double> csize;
  std::vector<std::vector<cd>> cc;           // cluster centres per level
  std::vector<std::vector<int>> cnear, cfar; // [m*9+slot], [m*27+slot]
  std::vector<cd> pos;
  std::vector<int> unkmap;
  std::vector<cd> nearb;                     // [9][64 test][64 basis]
  std::vector<cd> coeff;                     // [S_leaf][64]
  std::vector<cd> basis;                     // [64][S_leaf]
  std::vector<std::vector<double>> interp, anterp;
  std::vector<std::vector<int>> intind, antind;
  std::vector<std::vector<cd>> shm, shl, trans;
  std::vector<std::vector<int>> traid;
  std::vector<double> interph, anterph;
  std::vector<int> intindh, antindh;
  std::vector<std::vector<cd>> agm, agl;
  long matvec_in = 0, matvec_out = 0;
};

// ---------------------------------------------------------------- Green ----
// integrate.cpp:7-9
cd hn(int order, double x) { return cd(jn(order, x), yn(order, x)); }

// integrate.cpp:11-13: primitive of log(x^2+y^2) over a rectangle corner.
double integ_prim(double x, double y) {
  return -3 * x * y + x * y * std::log(x * x + y * y) + x * x * std::atan(y / x) +
         y * y * std::atan(x / y);
}

// integrate.cpp:15-41: (i/4) * pixel integral of H0^(1)(k0|r-r'|).
// Off the diagonal a midpoint rule; on the self term the small-argument
// model J0 + i(2/pi)(ln(k r/2) + gamma) integrated analytically.
cd green_pixel(const Oracle& o, cd pt, cd pb) {
  double d = std::abs(pt - pb);
  cd num(j0(o.k0 * d), 0.0), anal(0.0, 0.0);
  if (d < o.res / 8) {
    num = num + cd(0.0, 0.5772156649015329 * 2 / M_PI + 2 / M_PI * std::log(0.5));
    double xc = (pb - pt).real(), yc = (pb - pt).imag();
    double x0 = xc - o.res / 2, x1 = xc + o.res / 2;
    double y0v = yc - o.res / 2, y1 = yc + o.res / 2;
    double a = integ_prim(x1, y1) - integ_prim(x0, y1) - integ_prim(x1, y0v) + integ_prim(x0, y0v);
    anal = cd(0.0, (a / 2 + std::log(o.k0) * o.res * o.res) * 2 / M_PI);
  } else {
    num = num + cd(0.0, y0(o.k0 * d));
  }
  num = num * o.res;
  num = num * o.res;
  return (num + anal) * cd(0.0, 0.25);
}

// integrate.cpp:43-55 (P2M radiation coefficient of one pixel)
cd p2m_coeff(const Oracle& o, cd centre, cd pb, int k) {
  int S = o.nsamp[o.L - 1];
  double ang = 2 * M_PI * k / S;
  cd u(std::cos(ang), std::sin(ang));
  cd d = pb - centre;
  cd v = std::exp(cd(0.0, -o.k0 * (u.real() * d.real() + u.imag() * d.imag())));
  v = v * o.res;
  return v * o.res;
}

// integrate.cpp:57-69 (L2P receiving pattern of one pixel)
cd l2p_basis(const Oracle& o, cd centre, cd pt, int k) {
  int S = o.nsamp[o.L - 1];
  double ang = 2 * M_PI * k / S;
  cd u(std::cos(ang), std::sin(ang));
  cd d = pt - centre;
  return std::exp(cd(0.0, o.k0 * (u.real() * d.real() + u.imag() * d.imag())));
}

// Lagrange weights in the normalised angle coordinate, as used by the
// interpolators of setup_mlfma.cpp:573-599 and :652-677.
void lagrange_row(int m, int nfrom_out, int nto, int ninter, double* w, int* idx) {
  double ratio = (double)nto / nfrom_out;
  int centre = (ninter % 2 == 0) ? (int)std::ceil(m * ratio) : (int)std::round(m * ratio);
  double xm = (double)m / nfrom_out;
  for (int n = 0; n < ninter; n++) {
    int ind = centre + n - ninter / 2;
    double xn = (double)ind / nto;
    double mul = 1;
    for (int k = 0; k < ninter; k++)
      if (k != n) {
        double xk = (double)(centre + k - ninter / 2) / nto;
        mul = mul * (xm - xk) / (xn - xk);
      }
    w[n] = mul;
    ind = ind % nto;
    if (ind < 0) ind += nto;
    idx[n] = ind;
  }
}

// ------------------------------------------------------------- set-up ----
// setup_mlfma.cpp:70-223 (tree, centres, pixel positions, raster map, lists)
void build_tree(Oracle& o) {
  int L = o.L;
  o.nclus.assign(L, 0);
  o.csize.assign(L, 0);
  o.nterm.assign(L, 0);
  o.nsamp.assign(L, 0);
  for (int i = 0; i < L; i++) o.nclus[i] = 1 << (2 * i);
  for (int i = 0; i < L; i++) o.csize[i] = o.res * o.box * std::pow(2.0, L - i - 1);
  for (int i = 0; i < L; i++) {
    double kr = std::sqrt(2.0) * o.k0 * o.csize[i];
    double term = kr + 1.8 * std::pow((double)o.beta, 2.0 / 3) * std::pow(kr, 1.0 / 3);
    o.nterm[i] = (int)std::floor(term);
    o.nsamp[i] = 4 * o.nterm[i];
  }
  o.len = o.box << (L - 1);
  o.N = o.len * o.len;
  int bb = o.box * o.box;

  o.cc.assign(L, {});
  for (int i = 0; i < L; i++) o.cc[i].assign(o.nclus[i], cd(0, 0));
  o.pos.assign(o.N, cd(0, 0));
  o.cc[0][0] = cd(0, 0);
  for (int i = 0; i < L; i++) {
    for (int m = 0; m < o.nclus[i]; m++) {
      if (i < L - 1) {
        cd cor = o.cc[i][m] + o.csize[i] / 4 * cd(-1, 1);
        for (int k = 0; k < 2; k++)
          for (int l = 0; l < 2; l++)
            o.cc[i + 1][m * 4 + k * 2 + l] = cor + cd(l * o.csize[i + 1], -k * o.csize[i + 1]);
      } else {
        cd cor = o.cc[i][m] + (o.csize[i] / 2 - o.res / 2) * cd(-1, 1);
        for (int k = 0; k < o.box; k++)
          for (int l = 0; l < o.box; l++) o.pos[m * bb + k * o.box + l] = cor + cd(l * o.res, -k * o.res);
      }
    }
  }
  // raster map: setup_mlfma.cpp:177-185
  o.unkmap.assign(o.N, -1);
  cd cor = o.csize[0] / 2 * cd(-1, 1);
  for (int n = 0; n < o.N; n++) {
    int l = (int)std::floor((o.pos[n].real() - cor.real()) / o.res);
    int k = (int)std::floor((cor.imag() - o.pos[n].imag()) / o.res);
    o.unkmap[k * o.len + l] = n;
  }
  // near / far lists: setup_mlfma.cpp:187-223
  o.cnear.assign(L, {});
  o.cfar.assign(L, {});
  for (int i = 0; i < L; i++) {
    o.cnear[i].assign(o.nclus[i] * 9, -1);
    o.cfar[i].assign(o.nclus[i] * 27, -1);
  }
  o.cnear[0][0] = 0;
  for (int i = 1; i < L; i++) {
    double lim = o.csize[i] * std::sqrt(3.0);
#pragma omp parallel for
    for (int m = 0; m < o.nclus[i]; m++) {
      int nid = 0, fid = 0;
      int par = m / 4;
      for (int s = 0; s < 9; s++) {
        int uncle = o.cnear[i - 1][par * 9 + s];
        if (uncle == -1) continue;
        for (int c = 0; c < 4; c++) {
          int n = uncle * 4 + c;
          if (std::abs(o.cc[i][n] - o.cc[i][m]) < lim)
            o.cnear[i][m * 9 + nid++] = n;
          else
            o.cfar[i][m * 27 + fid++] = n;
        }
      }
    }
  }
  // leaf near lists re-sorted into geometric slots: setup_mlfma.cpp:459-484
  int Lf = L - 1;
  double a = o.csize[Lf];
  std::vector<int> resorted(o.nclus[Lf] * 9, -1);
  for (int m = 0; m < o.nclus[Lf]; m++)
    for (int s = 0; s < 9; s++) {
      int n = o.cnear[Lf][m * 9 + s];
      if (n == -1) continue;
      cd d = o.cc[Lf][m] - o.cc[Lf][n];
      int mi = (int)std::round(d.imag() / a) + 1;
      int ni = (int)std::round(-d.real() / a) + 1;
      resorted[m * 9 + mi * 3 + ni] = n;
    }
  o.cnear[Lf] = resorted;
}

// setup_mlfma.cpp:492-505: one 64x64 block per slot, taken from the first box
// (in serial order) that owns the slot.
void fill_near(Oracle& o) {
  int Lf = o.L - 1, bb = o.box * o.box;
  o.nearb.assign(9 * bb * bb, cd(0, 0));
  for (int s = 0; s < 9; s++) {
    int owner = -1, src = -1;
    for (int m = 0; m < o.nclus[Lf] && owner < 0; m++)
      if (o.cnear[Lf][m * 9 + s] != -1) {
        owner = m;
        src = o.cnear[Lf][m * 9 + s];
      }
    if (owner < 0) continue;
#pragma omp parallel for
    for (int e = 0; e < bb * bb; e++) {
      int k = e / bb, l = e % bb;
      o.nearb[s * bb * bb + k * bb + l] = green_pixel(o, o.pos[owner * bb + k], o.pos[src * bb + l]);
    }
  }
}

// setup_mlfma.cpp:506-543: diagonal translation operators on a 7x7 offset grid.
void fill_trans(Oracle& o) {
  int L = o.L;
  o.trans.assign(L, {});
  o.traid.assign(L, {});
  std::vector<cd> expn(o.nterm[0] + 1);
  for (int p = 0; p < o.nterm[0]; p++) expn[p] = std::exp(cd(0.0, (p + 1) * M_PI / 2));
  for (int i = 2; i < L; i++) {
    int S = o.nsamp[i];
    double a = o.csize[i];
    o.traid[i].assign(o.nclus[i] * 27, -1);
    o.trans[i].assign(49 * S, cd(0, 0));
    for (int m = 0; m < o.nclus[i]; m++)
      for (int s = 0; s < 27; s++) {
        int n = o.cfar[i][m * 27 + s];
        if (n == -1) continue;
        cd d = o.cc[i][n] - o.cc[i][m];
        int ni = (int)std::round(d.real() / a) + 3;
        int mi = (int)std::round(-d.imag() / a) + 3;
        o.traid[i][m * 27 + s] = mi * 7 + ni;
      }
    std::vector<cd> hank(o.nterm[i] + 1);
    for (int cm = 0; cm < 7; cm++)
      for (int cn = 0; cn < 7; cn++) {
        cd d = -cd(-3 * a + cn * a, 3 * a - cm * a);
        if (!(std::abs(d) > a * std::sqrt(3.0))) continue;
        for (int p = 0; p < o.nterm[i] + 1; p++) hank[p] = hn(p, o.k0 * std::abs(d));
        int base = cm * 7 * S + cn * S;
        double phi = std::arg(d);
#pragma omp parallel for
        for (int k = 0; k < S; k++) {
          cd red = hank[0];
          for (int p = 1; p < o.nterm[i] + 1; p++)
            red = red + 2 * std::cos(p * (2 * M_PI * k / S - phi)) * hank[p] * expn[p - 1];
          o.trans[i][base + k] = red;
        }
      }
  }
}

// setup_mlfma.cpp:544-553 (P2M / L2P), :554-570 (shifts), :571-631
// (interpolators and anterpolators), :642-695 (receiver interpolators).
void fill_tables(Oracle& o) {
  int L = o.L, bb = o.box * o.box, S = o.nsamp[L - 1], ni = o.ninter;
  o.coeff.assign(S * bb, cd(0, 0));
  o.basis.assign(bb * S, cd(0, 0));
  for (int n = 0; n < bb; n++)
    for (int k = 0; k < S; k++) {
      o.coeff[k * bb + n] = p2m_coeff(o, o.cc[L - 1][0], o.pos[n], k);
      o.basis[n * S + k] = l2p_basis(o, o.cc[L - 1][0], o.pos[n], k) / cd(S, 0) * cd(0.0, 0.25);
    }
  o.shm.assign(L, {});
  o.shl.assign(L, {});
  for (int i = 0; i < L - 1; i++) {
    int Si = o.nsamp[i];
    o.shm[i].assign(4 * Si, cd(0, 0));
    o.shl[i].assign(4 * Si, cd(0, 0));
    for (int c = 0; c < 4; c++) {
      cd rho = o.cc[i][0] - o.cc[i + 1][c];
      for (int k = 0; k < Si; k++) {
        double ang = 2 * M_PI * k / Si;
        cd kw = o.k0 * cd(std::cos(ang), std::sin(ang));
        double prod = kw.real() * rho.real() + kw.imag() * rho.imag();
        o.shm[i][c * Si + k] = std::exp(cd(0.0, prod));
        o.shl[i][c * Si + k] = std::exp(cd(0.0, -prod));
      }
    }
  }
  o.interp.assign(L, {});
  o.intind.assign(L, {});
  for (int i = 0; i < L - 1; i++) {
    o.interp[i].assign(o.nsamp[i] * ni, 0.0);
    o.intind[i].assign(o.nsamp[i] * ni, 0);
    for (int m = 0; m < o.nsamp[i]; m++)
      lagrange_row(m, o.nsamp[i], o.nsamp[i + 1], ni, &o.interp[i][m * ni], &o.intind[i][m * ni]);
  }
  o.anterp.assign(L, {});
  o.antind.assign(L, {});
  for (int i = 1; i < L; i++) {
    int Sc = o.nsamp[i], Sp = o.nsamp[i - 1];
    o.anterp[i].assign(Sc * 2 * ni, 0.0);
    o.antind[i].assign(Sc * 2 * ni, 0);
    std::vector<int> cnt(Sc, 0);
    for (int m = 0; m < Sp; m++)
      for (int n = 0; n < ni; n++) {
        int j = o.intind[i - 1][m * ni + n];
        if (cnt[j] >= 2 * ni) { fprintf(stderr, "oracle: anterpolator overflow\n"); continue; }
        o.antind[i][j * 2 * ni + cnt[j]] = m;
        o.anterp[i][j * 2 * ni + cnt[j]] = o.interp[i - 1][m * ni + n] * Sc / Sp;
        cnt[j]++;
      }
    // rotate each row so that its parent indices run contiguously (mod Sp)
    for (int j = 0; j < Sc; j++) {
      int* id = &o.antind[i][j * 2 * ni];
      double* w = &o.anterp[i][j * 2 * ni];
      for (int n = 1; n < cnt[j]; n++)
        if (id[n] != id[n - 1] + 1) {
          std::rotate(id, id + n, id + cnt[j]);
          std::rotate(w, w + n, w + cnt[j]);
          break;
        }
    }
  }
  // receivers
  int S0 = o.nsamp[0], R = o.numrx;
  o.interph.assign(R * ni, 0.0);
  o.intindh.assign(R * ni, 0);
  for (int m = 0; m < R; m++) lagrange_row(m, R, S0, ni, &o.interph[m * ni], &o.intindh[m * ni]);
  o.anterph.assign(S0 * 2 * ni, 0.0);
  o.antindh.assign(S0 * 2 * ni, 0);
  std::vector<int> cnt(S0, 0);
  for (int m = 0; m < R; m++)
    for (int n = 0; n < ni; n++) {
      int j = o.intindh[m * ni + n];
      if (cnt[j] >= 2 * ni) { fprintf(stderr, "oracle: receiver anterpolator overflow\n"); continue; }
      o.antindh[j * 2 * ni + cnt[j]] = m;
      o.anterph[j * 2 * ni + cnt[j]] = o.interph[m * ni + n] * S0 / R;
      cnt[j]++;
    }
  o.agm.assign(L, {});
  o.agl.assign(L, {});
  for (int i = 0; i < L; i++) {
    o.agm[i].assign((size_t)o.nclus[i] * o.nsamp[i], cd(0, 0));
    o.agl[i].assign((size_t)o.nclus[i] * o.nsamp[i], cd(0, 0));
  }
}

// ----------------------------------------------------------- operator ----
// solve.cpp:80-92 (P2M)
void p2m(Oracle& o, const cd* x) {
  int Lf = o.L - 1, S = o.nsamp[Lf], bb = o.box * o.box;
#pragma omp parallel for
  for (int b = 0; b < o.nclus[Lf]; b++)
    for (int k = 0; k < S; k++) {
      cd red = 0;
      for (int n = 0; n < bb; n++) red = red + o.coeff[k * bb + n] * x[b * bb + n];
      o.agm[Lf][(size_t)b * S + k] = red;
    }
}

// solve.cpp:93-112: interpolate each child to the parent's samples, shift, sum.
void m2m(Oracle& o, int i) {
  int Sp = o.nsamp[i], Sc = o.nsamp[i + 1], ni = o.ninter;
#pragma omp parallel for
  for (int m = 0; m < o.nclus[i]; m++)
    for (int k = 0; k < Sp; k++) {
      cd t1 = 0;
      for (int c = 0; c < 4; c++) {
        const cd* ch = &o.agm[i + 1][(size_t)(m * 4 + c) * Sc];
        cd red = 0;
        for (int t = 0; t < ni; t++) red = red + o.interp[i][k * ni + t] * ch[o.intind[i][k * ni + t]];
        t1 = t1 + red * o.shm[i][c * Sp + k];
      }
      o.agm[i][(size_t)m * Sp + k] = t1;
    }
}

// solve.cpp:133-148 (M2L, accumulating into zeroed locals)
void m2l(Oracle& o, int i) {
  int S = o.nsamp[i];
#pragma omp parallel for
  for (int m = 0; m < o.nclus[i]; m++)
    for (int s = 0; s < 27; s++) {
      int n = o.cfar[i][m * 27 + s];
      if (n == -1) continue;
      const cd* T = &o.trans[i][(size_t)o.traid[i][m * 27 + s] * S];
      const cd* src = &o.agm[i][(size_t)n * S];
      cd* dst = &o.agl[i][(size_t)m * S];
      for (int k = 0; k < S; k++) dst[k] = dst[k] + T[k] * src[k];
    }
}

// solve.cpp:149-169 (accumulate) and :388-429 (assign): shift the parent's
// locals toward each child, then anterpolate to the child's samples.
void l2l(Oracle& o, int i, bool accumulate) {
  int Sp = o.nsamp[i], Sc = o.nsamp[i + 1], na = 2 * o.ninter;
#pragma omp parallel
  {
    std::vector<cd> tmp(Sp);
#pragma omp for
    for (int n = 0; n < o.nclus[i]; n++)
      for (int c = 0; c < 4; c++) {
        for (int k = 0; k < Sp; k++) tmp[k] = o.shl[i][c * Sp + k] * o.agl[i][(size_t)n * Sp + k];
        cd* dst = &o.agl[i + 1][(size_t)(n * 4 + c) * Sc];
        for (int k = 0; k < Sc; k++) {
          cd red = 0;
          for (int t = 0; t < na; t++) red = red + o.anterp[i + 1][k * na + t] * tmp[o.antind[i + 1][k * na + t]];
          dst[k] = accumulate ? dst[k] + red : red;
        }
      }
  }
}

// solve.cpp:170-182 (L2P, assigns)
void l2p(Oracle& o, cd* r) {
  int Lf = o.L - 1, S = o.nsamp[Lf], bb = o.box * o.box;
#pragma omp parallel for
  for (int b = 0; b < o.nclus[Lf]; b++)
    for (int n = 0; n < bb; n++) {
      cd red = 0;
      for (int k = 0; k < S; k++) red = red + o.basis[n * S + k] * o.agl[Lf][(size_t)b * S + k];
      r[(size_t)b * bb + n] = red;
    }
}

// solve.cpp:199-216 (P2P, adds onto r)
void p2p(Oracle& o, const cd* x, cd* r) {
  int Lf = o.L - 1, bb = o.box * o.box;
#pragma omp parallel for
  for (int b = 0; b < o.nclus[Lf]; b++)
    for (int m = 0; m < bb; m++) {
      cd red = 0;
      for (int s = 0; s < 9; s++) {
        int nb = o.cnear[Lf][b * 9 + s];
        if (nb == -1) continue;
        const cd* blk = &o.nearb[(size_t)s * bb * bb + m * bb];
        for (int n = 0; n < bb; n++) red = red + x[(size_t)nb * bb + n] * blk[n];
      }
      r[(size_t)b * bb + m] = r[(size_t)b * bb + m] + red;
    }
}

// solve.cpp:25-248: r = G x.
void mlfma(Oracle& o, const cd* x, cd* r) {
  o.matvec_in++;
  for (int i = 0; i < o.L; i++) std::fill(o.agl[i].begin(), o.agl[i].end(), cd(0, 0));
  p2m(o, x);
  for (int i = o.L - 2; i > 1; i--) m2m(o, i);
  for (int i = 2; i < o.L; i++) m2l(o, i);
  for (int i = 2; i < o.L - 1; i++) l2l(o, i, true);
  l2p(o, r);
  p2p(o, x, r);
}

// solve.cpp:8-24: r = conj(G conj(x)); x is restored.
void mlfmah(Oracle& o, cd* x, cd* r) {
  for (int n = 0; n < o.N; n++) x[n] = std::conj(x[n]);
  mlfma(o, x, r);
  for (int n = 0; n < o.N; n++) {
    r[n] = std::conj(r[n]);
    x[n] = std::conj(x[n]);
  }
}

// Finite-radius mode: receiver m at radius (cos, sin)(2 pi m / numrx) and the
// midpoint-rule Green's function weight (i/4) H0(k0 |rho_m - r_n|) res^2.
cd circle_point(double radius, int m, int count) {
  double a = 2 * M_PI * m / count;
  return cd(radius * std::cos(a), radius * std::sin(a));
}
cd rx_weight(const Oracle& o, int m, int n) {
  double d = std::abs(circle_point(o.radius, m, o.numrx) - o.pos[n]);
  return cd(0, 0.25) * cd(j0(o.k0 * d), y0(o.k0 * d)) * (o.res * o.res);
}

// solve.cpp:249-355: far-field pattern at the receiver angles (radius 0), or
// the scattered field at the receiver circle (radius > 0, direct sums).
void aggregate(Oracle& o, const cd* x, cd* r) {
  if (o.radius > 0) {
#pragma omp parallel for
    for (int m = 0; m < o.numrx; m++) {
      cd s = 0;
      for (int n = 0; n < o.N; n++) s += rx_weight(o, m, n) * x[n];
      r[m] = s;
    }
    return;
  }
  o.matvec_out++;
  p2m(o, x);
  for (int i = o.L - 2; i > 1; i--) m2m(o, i);
  for (int i = 1; i > -1; i--)
    if (i + 1 < o.L) m2m(o, i);
  int ni = o.ninter;
  for (int m = 0; m < o.numrx; m++) {
    cd red = 0;
    for (int t = 0; t < ni; t++) red = red + o.interph[m * ni + t] * o.agm[0][o.intindh[m * ni + t]];
    r[m] = red;
  }
}

// solve.cpp:356-456: receiver back-projection; x is the output.  Radius > 0:
// the exact adjoint of the direct-sum aggregate.
void aggregateh(Oracle& o, cd* x, const cd* r) {
  if (o.radius > 0) {
#pragma omp parallel for
    for (int n = 0; n < o.N; n++) {
      cd s = 0;
      for (int m = 0; m < o.numrx; m++) s += std::conj(rx_weight(o, m, n)) * r[m];
      x[n] = s;
    }
    return;
  }
  o.matvec_out++;
  int na = 2 * o.ninter, S0 = o.nsamp[0];
  for (int m = 0; m < S0; m++) {
    cd red = 0;
    for (int t = 0; t < na; t++) red = red + o.anterph[m * na + t] * r[o.antindh[m * na + t]];
    o.agl[0][m] = red * cd(0, -4) * cd(o.numrx * o.res * o.res, 0);
  }
  for (int i = 0; i < 2 && i < o.L - 1; i++) l2l(o, i, false);
  for (int i = 2; i < o.L - 1; i++) l2l(o, i, false);
  l2p(o, x);
}

// Dense direct MoM product r = G x with G_mn = green_pixel(r_m, r_n): the
// matrix that the near blocks sample (setup_mlfma.cpp:503).  O(N^2).
void direct(Oracle& o, const cd* x, cd* r, int row0, int nrow) {
#pragma omp parallel for
  for (int m = row0; m < row0 + nrow; m++) {
    cd red = 0;
    for (int n = 0; n < o.N; n++) red = red + green_pixel(o, o.pos[m], o.pos[n]) * x[n];
    r[m - row0] = red;
  }
}

// -------------------------------------------------------------- solver ----
double norm2(const cd* a, int n) {
  double s = 0;
  for (int i = 0; i < n; i++) s += std::norm(a[i]);
  return s;
}
cd inner(const cd* a, const cd* b, int n) {
  cd s = 0;
  for (int i = 0; i < n; i++) s = s + std::conj(a[i]) * b[i];
  return s;
}

// bicgs_gpu.cu:61-76 + kernels.cu:394-409: forward v = x - G(o x),
// adjoint v = x - conj(o G(conj x)).
void apply_a(Oracle& o, const cd* x, const cd* g, cd* v, bool herm, std::vector<cd>& buf) {
  int N = o.N;
  if (herm) {
    for (int n = 0; n < N; n++) buf[n] = std::conj(x[n]);
    mlfma(o, buf.data(), v);
    for (int n = 0; n < N; n++) v[n] = x[n] - std::conj(g[n] * v[n]);
  } else {
    for (int n = 0; n < N; n++) buf[n] = x[n] * g[n];
    mlfma(o, buf.data(), v);
    for (int n = 0; n < N; n++) v[n] = x[n] - v[n];
  }
}

struct SolveInfo {
  int iters;
  int matvecs;
  double relres;
  int status;  // 0 converged, 1 not converged, 2 rho breakdown, 3 omega breakdown
};

// bicgs_gpu.cu:116-224 (unpreconditioned BiCGSTAB, x in/out).
SolveInfo bicgs(Oracle& o, cd* x, const cd* g, const cd* b, bool herm, double tol, int maxit) {
  int N = o.N;
  std::vector<cd> r(N), rt(N), p(N), v(N), s(N), t(N), buf(N), tmp(N);
  SolveInfo info{0, 0, 0.0, 1};
  double bn = std::sqrt(norm2(b, N));
  if (bn < 1e-20) bn = 1;
  apply_a(o, x, g, r.data(), herm, buf);
  info.matvecs++;
  for (int n = 0; n < N; n++) r[n] = b[n] + cd(-1, 0) * r[n];
  double err = std::sqrt(norm2(r.data(), N)) / bn;
  cd rho, rho1, alpha, omega, beta;
  if (err > tol) {
    rt = r;
    while (info.iters < maxit) {
      info.iters++;
      rho = inner(rt.data(), r.data(), N);
      if (std::abs(rho) < 1e-200) { info.status = 2; break; }
      if (info.iters == 1) {
        p = r;
      } else {
        beta = (rho / rho1) * (alpha / omega);
        for (int n = 0; n < N; n++) tmp[n] = p[n] + (-omega) * v[n];
        for (int n = 0; n < N; n++) p[n] = r[n] + beta * tmp[n];
      }
      apply_a(o, p.data(), g, v.data(), herm, buf);
      info.matvecs++;
      alpha = rho / inner(rt.data(), v.data(), N);
      for (int n = 0; n < N; n++) s[n] = r[n] + (-alpha) * v[n];
      apply_a(o, s.data(), g, t.data(), herm, buf);
      info.matvecs++;
      cd ts = inner(t.data(), s.data(), N);
      double tt = norm2(t.data(), N);
      omega = ts / tt;
      if (std::abs(ts) < 1e-200 && tt < 1e-200) omega = 1.0;
      for (int n = 0; n < N; n++) x[n] = x[n] + alpha * p[n];
      for (int n = 0; n < N; n++) x[n] = x[n] + omega * s[n];
      for (int n = 0; n < N; n++) r[n] = s[n] + (-omega) * t[n];
      err = std::sqrt(norm2(r.data(), N)) / bn;
      if (err < tol) break;
      if (std::abs(omega) < 1e-200) { info.status = 3; break; }
      rho1 = rho;
    }
  }
  info.relres = err;
  if (err < tol) info.status = 0;
  return info;
}

// ---------------------------------------------------------------- DBIM ----
struct Dbim {
  int numtx = 0;
  double regular = 0;
  double tol = 1e-4;
  int maxit = 300;
  std::vector<cd> tx;        // transmitter directions (mom.cpp:141-147)
  std::vector<cd> inc, tot;  // [numtx][N]
  std::vector<cd> mes, mesd, rhs;  // [numtx][numrx]
  std::vector<cd> o, od, ot, save, omes;
  double mesnorm = 0, objnorm = 0;
  int iter = 0;
  long solves = 0, solve_iters = 0;
  double hist[5] = {0, 0, 0, 0, 0};
  // planted mistakes for the tests that pin this loop (1 = the reference's
  // algorithm): scale of alpha, of beta and of the regulariser's gradient term
  double mut_alpha = 1, mut_beta = 1, mut_reg = 1;
};

void track(Dbim& d, const SolveInfo& si) {
  d.solves++;
  d.solve_iters += si.iters;
}

// mom.cpp:247-271: gradient of the misfit summed over transmitters.
void dbim_gradient(Oracle& o, Dbim& d) {
  int N = o.N, R = o.numrx;
  double kk = o.k0;
  std::vector<cd> buff(N, cd(0, 0)), y(N), c(N), g(N), x(N);
  for (int i = 0; i < d.numtx; i++) {
    aggregateh(o, y.data(), &d.mesd[(size_t)i * R]);
    for (int n = 0; n < N; n++) c[n] = std::conj(d.o[n] * kk * kk) * y[n];
    std::fill(x.begin(), x.end(), cd(0, 0));
    for (int n = 0; n < N; n++) g[n] = d.o[n] * kk * kk;
    track(d, bicgs(o, x.data(), g.data(), c.data(), true, d.tol, d.maxit));
    mlfmah(o, x.data(), c.data());
    for (int n = 0; n < N; n++) y[n] = y[n] + c[n];
    for (int n = 0; n < N; n++) buff[n] = buff[n] + std::conj(d.tot[(size_t)i * N + n]) * y[n];
    for (int n = 0; n < N; n++) buff[n] = buff[n] + d.mut_reg * d.regular * d.regular * d.o[n];
  }
  d.od = buff;
}

// setup_born.cpp:171-242 (metrics; the cg_<it>.bin write is left to callers)
void dbim_error(Oracle& o, Dbim& d) {
  double err = 0, errd = 0, obj = 0, objn = 0, errm = 0;
  for (size_t m = 0; m < d.mesd.size(); m++) err += std::norm(d.mesd[m]);
  for (int n = 0; n < o.N; n++) errd += std::norm(d.od[n]);
  for (int n = 0; n < o.N; n++) {
    obj += std::norm(d.regular * d.o[n]);
    objn += std::norm(d.o[n]);
    errm += std::norm(d.o[n] - d.omes[n]);
  }
  d.hist[0] = err / d.mesnorm;
  d.hist[1] = errm / d.objnorm;
  d.hist[2] = obj / d.mesnorm;
  d.hist[3] = errd;
  d.hist[4] = objn;
}

// setup_born.cpp:62-65, :139-168 and mom.cpp:222-276: incident fields,
// synthetic measurements from the true contrast, zero initial object,
// first gradient.
void dbim_setup(Oracle& o, Dbim& d, const cd* otrue) {
  int N = o.N, R = o.numrx, T = d.numtx;
  double kk = o.k0;
  d.tx.resize(T);
  for (int i = 0; i < T; i++) {
    double ang = 360.0 / T * i;
    d.tx[i] = cd(std::cos(ang / 360 * 2 * M_PI), std::sin(ang / 360 * 2 * M_PI));
  }
  d.inc.assign((size_t)T * N, cd(0, 0));
  for (int i = 0; i < T; i++)
    for (int n = 0; n < N; n++) {
      if (o.radius > 0) {  // line source at the transmitter circle: (i/4) H0(k0 |r - rho_t|)
        double dd = std::abs(o.pos[n] - circle_point(o.radius, i, T));
        d.inc[(size_t)i * N + n] = cd(0, 0.25) * cd(j0(kk * dd), y0(kk * dd));
      } else {
        d.inc[(size_t)i * N + n] =
            std::exp(cd(0, kk * (d.tx[i].real() * o.pos[n].real() + d.tx[i].imag() * o.pos[n].imag())));
      }
    }
  d.o.assign(otrue, otrue + N);
  d.mes.assign((size_t)T * R, cd(0, 0));
  std::vector<cd> x(N), g(N);
  for (int i = 0; i < T; i++) {
    std::fill(x.begin(), x.end(), cd(0, 0));
    for (int n = 0; n < N; n++) g[n] = d.o[n] * kk * kk;
    track(d, bicgs(o, x.data(), g.data(), &d.inc[(size_t)i * N], false, d.tol, d.maxit));
    for (int n = 0; n < N; n++) x[n] = x[n] * g[n];
    aggregate(o, x.data(), &d.mes[(size_t)i * R]);
  }
  d.omes = d.o;
  d.objnorm = 0;
  for (int n = 0; n < N; n++) d.objnorm += std::norm(d.omes[n]);
  d.mesnorm = 0;
  for (size_t m = 0; m < d.mes.size(); m++) d.mesnorm += std::norm(d.mes[m]);
  // mom.cpp:224-229: initial guess (no object)
  std::fill(d.o.begin(), d.o.end(), cd(0, 0));
  d.tot = d.inc;
  d.mesd.resize(d.mes.size());
  for (size_t m = 0; m < d.mes.size(); m++) d.mesd[m] = -d.mes[m];
  d.rhs.assign(d.mes.size(), cd(0, 0));
  dbim_gradient(o, d);
  dbim_error(o, d);
  d.ot = d.od;
  d.iter = 0;
}

// mom.cpp:277-400: one DBIM / nonlinear-CG iteration.
double dbim_iterate(Oracle& o, Dbim& d) {
  int N = o.N, R = o.numrx, T = d.numtx;
  double kk = o.k0;
  std::vector<cd> y(N), c(N), g(N), x(N);
  // step length (mom.cpp:280-316)
  for (int i = 0; i < T; i++) {
    for (int n = 0; n < N; n++) y[n] = d.tot[(size_t)i * N + n] * d.ot[n] * kk * kk;
    mlfma(o, y.data(), c.data());
    std::fill(x.begin(), x.end(), cd(0, 0));
    for (int n = 0; n < N; n++) g[n] = d.o[n] * kk * kk;
    track(d, bicgs(o, x.data(), g.data(), c.data(), false, d.tol, d.maxit));
    for (int n = 0; n < N; n++) y[n] = y[n] + d.o[n] * kk * kk * x[n];
    aggregate(o, y.data(), &d.rhs[(size_t)i * R]);
  }
  cd t1 = 0;
  double t2 = 0;
  for (int i = 0; i < T; i++)
    for (int m = 0; m < R; m++) {
      t1 = t1 + std::conj(d.mesd[(size_t)i * R + m]) * d.rhs[(size_t)i * R + m];
      t2 = t2 + std::norm(d.rhs[(size_t)i * R + m]);
    }
  double alpha = d.mut_alpha * t1.real() / t2;
  for (int n = 0; n < N; n++) d.o[n] = d.o[n] - alpha * d.ot[n];
  // total-field update (mom.cpp:322-336)
  for (int i = 0; i < T; i++) {
    cd* tot = &d.tot[(size_t)i * N];
    std::fill(tot, tot + N, cd(0, 0));
    for (int n = 0; n < N; n++) g[n] = d.o[n] * kk * kk;
    track(d, bicgs(o, tot, g.data(), &d.inc[(size_t)i * N], false, d.tol, d.maxit));
    for (int n = 0; n < N; n++) x[n] = tot[n] * d.o[n] * kk * kk;
    aggregate(o, x.data(), &d.rhs[(size_t)i * R]);
    for (int m = 0; m < R; m++) d.mesd[(size_t)i * R + m] = d.rhs[(size_t)i * R + m] - d.mes[(size_t)i * R + m];
  }
  d.save = d.od;
  dbim_gradient(o, d);
  dbim_error(o, d);
  // Polak-Ribiere direction (mom.cpp:383-399)
  cd n1 = 0;
  double n2 = 0;
  for (int n = 0; n < N; n++) {
    n1 = n1 + d.od[n] * std::conj(d.od[n] - d.save[n]);
    n2 = n2 + std::norm(d.save[n]);
  }
  cd beta = d.mut_beta * n1 / n2;
  for (int n = 0; n < N; n++) d.ot[n] = d.od[n] + beta * d.ot[n];
  d.iter++;
  return alpha;
}

}  // namespace

// ================================================================ C ABI ====
// Complex arrays are interleaved (re, im) doubles, i.e. std::complex<double>.
// copy a table into a caller's array (nothing for an empty table, whose data()
// may be null)
template <typename T, typename U>
void copy_out(U* out, const std::vector<T>& v) {
  if (!v.empty()) memcpy(out, v.data(), sizeof(T) * v.size());
}

extern "C" {

void* orc_create(int level, double res, double scale, int beta, int ninter, int numrx) {
  Oracle* o = new Oracle();
  o->L = level;
  o->res = res;
  o->k0 = 2 * M_PI * scale;
  o->beta = beta;
  o->ninter = ninter;
  o->numrx = numrx;
  build_tree(*o);
  fill_near(*o);
  fill_trans(*o);
  fill_tables(*o);
  return o;
}
void orc_destroy(void* h) { delete (Oracle*)h; }
void orc_set_radius(void* h, double radius) { ((Oracle*)h)->radius = radius; }

int orc_numunk(void* h) { return ((Oracle*)h)->N; }
int orc_numsamp(void* h, int i) { return ((Oracle*)h)->nsamp[i]; }
int orc_numterm(void* h, int i) { return ((Oracle*)h)->nterm[i]; }
double orc_k0(void* h) { return ((Oracle*)h)->k0; }
double orc_clusize(void* h, int i) { return ((Oracle*)h)->csize[i]; }

void orc_pos(void* h, double* out) {
  Oracle* o = (Oracle*)h;
  memcpy(out, o->pos.data(), sizeof(cd) * o->N);
}
void orc_unkmap(void* h, int* out) {
  Oracle* o = (Oracle*)h;
  memcpy(out, o->unkmap.data(), sizeof(int) * o->N);
}
void orc_centres(void* h, int i, double* out) {
  Oracle* o = (Oracle*)h;
  memcpy(out, o->cc[i].data(), sizeof(cd) * o->nclus[i]);
}
void orc_near_list(void* h, int i, int* out) {
  Oracle* o = (Oracle*)h;
  copy_out(out, o->cnear[i]);
}
void orc_far_list(void* h, int i, int* out) {
  Oracle* o = (Oracle*)h;
  copy_out(out, o->cfar[i]);
}
void orc_traid(void* h, int i, int* out) {
  Oracle* o = (Oracle*)h;
  copy_out(out, o->traid[i]);
}
void orc_near_blocks(void* h, double* out) {
  Oracle* o = (Oracle*)h;
  copy_out(out, o->nearb);
}
void orc_coeff(void* h, double* out) {
  Oracle* o = (Oracle*)h;
  copy_out(out, o->coeff);
}
void orc_basis(void* h, double* out) {
  Oracle* o = (Oracle*)h;
  copy_out(out, o->basis);
}
void orc_trans(void* h, int i, double* out) {
  Oracle* o = (Oracle*)h;
  copy_out(out, o->trans[i]);
}
void orc_shift(void* h, int i, double* mul, double* loc) {
  Oracle* o = (Oracle*)h;
  copy_out(mul, o->shm[i]);
  copy_out(loc, o->shl[i]);
}
void orc_interp(void* h, int i, double* w, int* idx) {
  Oracle* o = (Oracle*)h;
  copy_out(w, o->interp[i]);
  copy_out(idx, o->intind[i]);
}
void orc_anterp(void* h, int i, double* w, int* idx) {
  Oracle* o = (Oracle*)h;
  copy_out(w, o->anterp[i]);
  copy_out(idx, o->antind[i]);
}
void orc_rx_interp(void* h, double* w, int* idx, double* aw, int* aidx) {
  Oracle* o = (Oracle*)h;
  copy_out(w, o->interph);
  copy_out(idx, o->intindh);
  copy_out(aw, o->anterph);
  copy_out(aidx, o->antindh);
}
// multipoles (which=0) or locals (which=1) of level i after the last operator call
void orc_stage(void* h, int which, int i, double* out) {
  Oracle* o = (Oracle*)h;
  copy_out(out, which == 0 ? o->agm[i] : o->agl[i]);
}

void orc_green(void* h, double tr, double ti, double br, double bi, double* out) {
  cd v = green_pixel(*(Oracle*)h, cd(tr, ti), cd(br, bi));
  out[0] = v.real();
  out[1] = v.imag();
}
void orc_hn(int order, double x, double* out) {
  cd v = hn(order, x);
  out[0] = v.real();
  out[1] = v.imag();
}

void orc_mlfma(void* h, const double* x, double* r) { mlfma(*(Oracle*)h, (const cd*)x, (cd*)r); }
void orc_mlfmah(void* h, double* x, double* r) { mlfmah(*(Oracle*)h, (cd*)x, (cd*)r); }
void orc_aggregate(void* h, const double* x, double* r) { aggregate(*(Oracle*)h, (const cd*)x, (cd*)r); }
void orc_aggregateh(void* h, double* x, const double* r) { aggregateh(*(Oracle*)h, (cd*)x, (const cd*)r); }
void orc_direct(void* h, const double* x, double* r, int row0, int nrow) {
  direct(*(Oracle*)h, (const cd*)x, (cd*)r, row0, nrow);
}
// per-stage entry points for stage parity
void orc_p2p_only(void* h, const double* x, double* r) {
  Oracle* o = (Oracle*)h;
  std::fill((cd*)r, (cd*)r + o->N, cd(0, 0));
  p2p(*o, (const cd*)x, (cd*)r);
}

// returns status; info = {iters, matvecs, relres}
int orc_bicgs(void* h, double* x, const double* g, const double* b, int herm, double tol, int maxit,
              double* info) {
  SolveInfo si = bicgs(*(Oracle*)h, (cd*)x, (const cd*)g, (const cd*)b, herm != 0, tol, maxit);
  info[0] = si.iters;
  info[1] = si.matvecs;
  info[2] = si.relres;
  return si.status;
}

void* orc_dbim_create(void* h, int numtx, double regular, double tol, int maxit, const double* otrue) {
  Oracle* o = (Oracle*)h;
  Dbim* d = new Dbim();
  d->numtx = numtx;
  d->regular = regular;
  d->tol = tol;
  d->maxit = maxit;
  dbim_setup(*o, *d, (const cd*)otrue);
  return d;
}
void orc_dbim_destroy(void* d) { delete (Dbim*)d; }
// test hook: plant a scale error in alpha, beta or the regulariser term
void orc_dbim_mutate(void* dd, double alpha_scale, double beta_scale, double reg_scale) {
  Dbim* d = (Dbim*)dd;
  d->mut_alpha = alpha_scale;
  d->mut_beta = beta_scale;
  d->mut_reg = reg_scale;
}
double orc_dbim_iterate(void* h, void* dd) { return dbim_iterate(*(Oracle*)h, *(Dbim*)dd); }
void orc_dbim_hist(void* dd, double* out) { memcpy(out, ((Dbim*)dd)->hist, sizeof(double) * 5); }
void orc_dbim_get(void* h, void* dd, int which, double* out) {
  Oracle* o = (Oracle*)h;
  Dbim* d = (Dbim*)dd;
  const std::vector<cd>* v = nullptr;
  switch (which) {
    case 0: v = &d->o; break;
    case 1: v = &d->od; break;
    case 2: v = &d->ot; break;
    case 3: v = &d->mes; break;
    case 4: v = &d->mesd; break;
    case 5: v = &d->inc; break;
    case 6: v = &d->tot; break;
    default: return;
  }
  (void)o;
  copy_out(out, *v);
}
void orc_dbim_counts(void* dd, long* out) {
  Dbim* d = (Dbim*)dd;
  out[0] = d->solves;
  out[1] = d->solve_iters;
}
long orc_matvecs(void* h) { return ((Oracle*)h)->matvec_in; }
int orc_threads() { return omp_get_max_threads(); }

}  // extern "C"
