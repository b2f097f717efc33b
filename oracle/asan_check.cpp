// asan_check.cpp -- host AddressSanitizer / UndefinedBehaviorSanitizer run of
// the CPU oracle (test infrastructure only; built by `make -C oracle asan`).
//
// Exercises every table accessor and operator of isg_oracle.cpp at 32x32 and
// 64x64 pixels (levels 3, 4), plane-wave and finite-radius geometry: the
// tables, mlfma / mlfmah / aggregate / aggregateh / the direct product,
// BiCGSTAB forward and adjoint, and two DBIM iterations.  Any out-of-bounds
// access, use after free or undefined arithmetic aborts the run; a clean run
// prints ASAN CHECK OK.
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <vector>

extern "C" {
void* orc_create(int level, double res, double scale, int beta, int ninter, int numrx);
void orc_destroy(void* h);
void orc_set_radius(void* h, double radius);
int orc_numunk(void* h);
int orc_numsamp(void* h, int i);
double orc_k0(void* h);
void orc_pos(void* h, double* out);
void orc_unkmap(void* h, int* out);
void orc_centres(void* h, int i, double* out);
void orc_near_list(void* h, int i, int* out);
void orc_far_list(void* h, int i, int* out);
void orc_traid(void* h, int i, int* out);
void orc_near_blocks(void* h, double* out);
void orc_coeff(void* h, double* out);
void orc_basis(void* h, double* out);
void orc_trans(void* h, int i, double* out);
void orc_shift(void* h, int i, double* mul, double* loc);
void orc_interp(void* h, int i, double* w, int* idx);
void orc_anterp(void* h, int i, double* w, int* idx);
void orc_rx_interp(void* h, double* w, int* idx, double* aw, int* aidx);
void orc_stage(void* h, int which, int i, double* out);
void orc_mlfma(void* h, const double* x, double* r);
void orc_mlfmah(void* h, double* x, double* r);
void orc_aggregate(void* h, const double* x, double* r);
void orc_aggregateh(void* h, double* x, const double* r);
void orc_direct(void* h, const double* x, double* r, int row0, int nrow);
void orc_p2p_only(void* h, const double* x, double* r);
int orc_bicgs(void* h, double* x, const double* g, const double* b, int herm, double tol, int maxit, double* info);
void* orc_dbim_create(void* h, int numtx, double regular, double tol, int maxit, const double* otrue);
void orc_dbim_destroy(void* d);
double orc_dbim_iterate(void* h, void* dd);
void orc_dbim_hist(void* dd, double* out);
}

typedef std::complex<double> cd;

static void run(int level, double radius) {
  const int numrx = 16, ninter = 30;
  void* h = orc_create(level, 0.1, 1.0, 5, ninter, numrx);
  if (radius > 0) orc_set_radius(h, radius);
  const int N = orc_numunk(h);
  const double k0 = orc_k0(h);
  std::vector<double> pos(2 * N), buf, buf2;
  std::vector<int> ibuf, ibuf2;
  orc_pos(h, pos.data());
  ibuf.resize(N);
  orc_unkmap(h, ibuf.data());
  for (int i = 0; i < level; i++) {
    const int nc = 1 << (2 * i), S = orc_numsamp(h, i);
    buf.resize(2 * nc);
    orc_centres(h, i, buf.data());
    ibuf.resize(nc * 27);
    orc_near_list(h, i, ibuf.data());
    orc_far_list(h, i, ibuf.data());
    orc_traid(h, i, ibuf.data());
    if (i >= 2) {
      buf.resize(2 * 49 * S);
      orc_trans(h, i, buf.data());
    }
    buf.resize(8 * S);
    buf2.resize(8 * S);
    if (i < level - 1) orc_shift(h, i, buf.data(), buf2.data());
    buf.resize(2 * ninter * S);
    ibuf.resize(2 * ninter * S);
    if (i < level - 1) orc_interp(h, i, buf.data(), ibuf.data());
    if (i > 0) orc_anterp(h, i, buf.data(), ibuf.data());
  }
  const int S0 = orc_numsamp(h, 0), Sl = orc_numsamp(h, level - 1);
  buf.resize(2 * 9 * 64 * 64);
  orc_near_blocks(h, buf.data());
  buf.resize(2 * 64 * Sl);
  orc_coeff(h, buf.data());
  orc_basis(h, buf.data());
  buf.resize(numrx * ninter);
  ibuf.resize(numrx * ninter);
  buf2.resize(S0 * 2 * ninter);
  ibuf2.resize(S0 * 2 * ninter);
  orc_rx_interp(h, buf.data(), ibuf.data(), buf2.data(), ibuf2.data());

  std::vector<cd> x(N), r(N), g(N), b(N), rr(numrx);
  srand(7);
  for (int n = 0; n < N; n++) {
    x[n] = cd(rand() / (double)RAND_MAX - 0.5, rand() / (double)RAND_MAX - 0.5);
    const cd p(pos[2 * n], pos[2 * n + 1]);
    g[n] = 0.05 * std::exp(-std::norm(p) / 2.0) * k0 * k0;
    b[n] = std::exp(cd(0, k0 * p.real()));
  }
  orc_mlfma(h, (double*)x.data(), (double*)r.data());
  for (int i = 0; i < level; i++) {
    buf.resize(2 * (1 << (2 * i)) * orc_numsamp(h, i));
    orc_stage(h, 0, i, buf.data());
    orc_stage(h, 1, i, buf.data());
  }
  orc_mlfmah(h, (double*)x.data(), (double*)r.data());
  orc_aggregate(h, (double*)x.data(), (double*)rr.data());
  orc_aggregateh(h, (double*)r.data(), (double*)rr.data());
  orc_direct(h, (double*)x.data(), (double*)r.data(), 0, 64);
  orc_p2p_only(h, (double*)x.data(), (double*)r.data());
  double info[3];
  for (int herm = 0; herm < 2; herm++) {
    std::vector<cd> s(N, cd(0, 0));
    orc_bicgs(h, (double*)s.data(), (double*)g.data(), (double*)b.data(), herm, 1e-8, 300, info);
  }
  std::vector<cd> ot(N);
  for (int n = 0; n < N; n++) ot[n] = g[n] / (k0 * k0);
  void* d = orc_dbim_create(h, 4, 0.03, 1e-6, 300, (double*)ot.data());
  double hist[5];
  for (int it = 0; it < 2; it++) orc_dbim_iterate(h, d);
  orc_dbim_hist(d, hist);
  orc_dbim_destroy(d);
  orc_destroy(h);
  printf("level %d radius %g: misfit %.3e\n", level, radius, hist[0]);
}

int main() {
  run(3, 0.0);
  run(4, 0.0);
  run(3, 5.0);
  printf("ASAN CHECK OK\n");
  return 0;
}
