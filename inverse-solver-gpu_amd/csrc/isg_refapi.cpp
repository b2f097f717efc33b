// isg_refapi.cpp -- the reference's C++ call surface over libisg.so.
//
// Compile this file together with the reference's driver (mom.cpp,
// setup_born.cpp) in place of setup_mlfma.cpp, gpu.cu, bicgs_gpu.cu,
// matvec_gpu.cu and kernels.cu, and link -lisg.  It defines the functions the
// driver calls with their reference signatures (vars.h:14-26) and reads the
// configuration from the driver's globals exactly where the reference does:
//   res, box, level, k0, numunk, numrx, toli, iti   (mom.cpp:4-22)
//   pos, unkmap                                      (allocated and filled by
//                                                      setup_mlfma, setup_mlfma.cpp:105,122)
// One process-global context backs every call (the reference keeps its state
// in globals too).  ISG_PRECISION=1 in the environment selects the mixed
// FP32-operator path for bicgs.  Errors abort with a message (the reference
// has no error returns).
#include <complex>
#include <cstdio>
#include <cstdlib>

#include "isg.h"

using std::complex;

extern double res;
extern int box;
extern int level;
extern double k0;
extern int numunk;
extern int numrx;
extern double toli;
extern int iti;
extern complex<double>* pos;
extern int* unkmap;

namespace {

isg_context* g_ctx = nullptr;

void die(const char* what, int rc) {
  fprintf(stderr, "isg: %s failed (%d): %s\n", what, rc, g_ctx ? isg_last_error(g_ctx) : "");
  std::exit(1);
}

}  // namespace

// setup_mlfma.cpp:50-640: operator set-up; also fills pos and unkmap.
void setup_mlfma() {
  isg_config c;
  isg_config_default(&c);
  c.level = level;
  c.box = box;
  c.res = res;
  c.scale = k0 / (2 * M_PI);
  c.numrx = numrx;
  c.toli = toli;
  c.iti = iti;
  const char* p = getenv("ISG_PRECISION");
  c.precision = p ? atoi(p) : 0;
  int rc = isg_create(&c, &g_ctx);
  if (rc != ISG_OK) die("isg_create", rc);
  pos = new complex<double>[numunk];
  unkmap = new int[numunk];
  rc = isg_geometry(g_ctx, reinterpret_cast<double*>(pos), unkmap);
  if (rc != ISG_OK) die("isg_geometry", rc);
}

// gpu.cu:69 and bicgs_gpu.cu:89: device set-up happens in isg_create.
void setup_gpu() {}
void setup_bicgs() {}

void mlfma(complex<double>* x, complex<double>* r) {
  int rc = isg_mlfma(g_ctx, reinterpret_cast<const double*>(x), reinterpret_cast<double*>(r));
  if (rc != ISG_OK) die("isg_mlfma", rc);
}

void mlfmah(complex<double>* x, complex<double>* r) {
  int rc = isg_mlfmah(g_ctx, reinterpret_cast<const double*>(x), reinterpret_cast<double*>(r));
  if (rc != ISG_OK) die("isg_mlfmah", rc);
}

void aggregate(complex<double>* x, complex<double>* r) {
  int rc = isg_aggregate(g_ctx, reinterpret_cast<const double*>(x), reinterpret_cast<double*>(r));
  if (rc != ISG_OK) die("isg_aggregate", rc);
}

void aggregateh(complex<double>* x, complex<double>* r) {
  int rc = isg_aggregateh(g_ctx, reinterpret_cast<double*>(x), reinterpret_cast<const double*>(r));
  if (rc != ISG_OK) die("isg_aggregateh", rc);
}

// bicgs_gpu.cu:116-224: prints what the reference prints -- non-convergence,
// then the iteration count and the solve's matrix-vector products
// (bicgs_gpu.cu:219-220).
void bicgs(complex<double>* x, complex<double>* o, complex<double>* b, bool herm) {
  isg_solve_info info;
  long long c0[5], c1[5];
  isg_counters(g_ctx, c0);
  int rc = isg_bicgs(g_ctx, reinterpret_cast<double*>(x), reinterpret_cast<const double*>(o),
                     reinterpret_cast<const double*>(b), herm ? 1 : 0, &info);
  if (rc < 0) die("isg_bicgs", rc);
  isg_counters(g_ctx, c1);
  if (info.status != ISG_SOLVE_CONVERGED)
    printf("NOT CONVERGED! status %d iter: %d relres %e\n", info.status, info.iters, info.relres);
  printf("NUMBER OF %d ITERATIONS: %d (%lld)\n", herm ? 1 : 0, info.iters, c1[0] - c0[0]);
}
