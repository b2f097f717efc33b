// refapi_driver.cpp -- a driver in the reference's style for the
// reference-named shim (inverse-solver-gpu_amd/csrc/isg_refapi.cpp).
//
// Like mom.cpp it defines the configuration globals (mom.cpp:4-22), leaves
// pos/unkmap to setup_mlfma (setup_mlfma.cpp:105,122), calls
// setup_mlfma/setup_gpu/setup_bicgs (mom.cpp:184-190) and then the per-
// transmitter calls of mom.cpp:250-350 and setup_born.cpp:144,148 on arrays
// it owns.  Inputs come from files a test writes, outputs go back as files:
//
//   refapi_driver DIR LEVEL NUMRX TOLI
//     DIR/x.bin  numunk complex   (operator input)
//     DIR/r.bin  numrx complex    (receiver data for aggregateh)
//     DIR/g.bin  numunk complex   (k0^2 * contrast, as mom.cpp:256 passes it)
//     DIR/b.bin  numunk complex   (right-hand side)
//   writes pos, unkmap, mlfma, mlfmah, aggregate, aggregateh, bicgs_f, bicgs_h
//   (.bin, raw little-endian) and prints the shim's BiCGSTAB lines.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

using std::complex;

// the reference driver's globals (mom.cpp:4-22)
double res = 0.1;
int box = 8;
int level = 4;
double k0 = 2 * M_PI;
int numunk = 0;
int numrx = 16;
double toli = 1e-4;
int iti = 300;
complex<double>* pos = nullptr;
int* unkmap = nullptr;

// the reference's call surface (vars.h:14-26)
void setup_mlfma();
void setup_gpu();
void setup_bicgs();
void mlfma(complex<double>* x, complex<double>* r);
void mlfmah(complex<double>* x, complex<double>* r);
void aggregate(complex<double>* x, complex<double>* r);
void aggregateh(complex<double>* x, complex<double>* r);
void bicgs(complex<double>* x, complex<double>* o, complex<double>* b, bool herm);

static std::string dir;

static void rd(const char* name, void* p, size_t bytes) {
  FILE* f = fopen((dir + "/" + name).c_str(), "rb");
  if (!f || fread(p, 1, bytes, f) != bytes) {
    fprintf(stderr, "cannot read %s\n", name);
    exit(2);
  }
  fclose(f);
}

static void wr(const char* name, const void* p, size_t bytes) {
  FILE* f = fopen((dir + "/" + name).c_str(), "wb");
  if (!f || fwrite(p, 1, bytes, f) != bytes) {
    fprintf(stderr, "cannot write %s\n", name);
    exit(2);
  }
  fclose(f);
}

int main(int argc, char** argv) {
  if (argc != 5) {
    fprintf(stderr, "usage: %s DIR LEVEL NUMRX TOLI\n", argv[0]);
    return 2;
  }
  dir = argv[1];
  level = atoi(argv[2]);
  numrx = atoi(argv[3]);
  toli = atof(argv[4]);
  const int side = box << (level - 1);
  numunk = side * side;  // mom.cpp:137-138
  const size_t vb = sizeof(complex<double>) * numunk, rb = sizeof(complex<double>) * numrx;

  setup_mlfma();
  setup_gpu();
  setup_bicgs();
  wr("pos.bin", pos, vb);
  wr("unkmap.bin", unkmap, sizeof(int) * numunk);

  std::vector<complex<double>> x(numunk), r(numrx), g(numunk), b(numunk), y(numunk), ry(numrx);
  rd("x.bin", x.data(), vb);
  rd("r.bin", r.data(), rb);
  rd("g.bin", g.data(), vb);
  rd("b.bin", b.data(), vb);

  mlfma(x.data(), y.data());
  wr("mlfma.bin", y.data(), vb);
  std::vector<complex<double>> x0 = x;
  mlfmah(x.data(), y.data());  // the reference conjugates x in place and restores it
  wr("mlfmah.bin", y.data(), vb);
  if (std::memcmp(x0.data(), x.data(), vb) != 0) printf("MLFMAH CHANGED ITS INPUT\n");
  aggregate(x.data(), ry.data());
  wr("aggregate.bin", ry.data(), rb);
  aggregateh(y.data(), r.data());  // output first (matvec_gpu.cu:369)
  wr("aggregateh.bin", y.data(), vb);

  // forward and adjoint solves from zero, as mom.cpp:254-258 and :286 call them
  std::fill(y.begin(), y.end(), complex<double>(0, 0));
  bicgs(y.data(), g.data(), b.data(), false);
  wr("bicgs_f.bin", y.data(), vb);
  std::fill(y.begin(), y.end(), complex<double>(0, 0));
  bicgs(y.data(), g.data(), b.data(), true);
  wr("bicgs_h.bin", y.data(), vb);
  printf("DRIVER DONE\n");
  return 0;
}
