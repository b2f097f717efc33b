// isg_plan.hpp -- host-side MLFMA operator plan for the MI355X solver.
//
// Builds the quadtree geometry and every operator table of the 2-D MLFMA
// (near-field blocks, radiation/receiving patterns, shifts, Lagrange
// interpolators and anterpolators, diagonal translation operators, receiver
// interpolators) directly in the layouts the HIP kernels read.  The
// mathematics is that of the reference's setup (setup_mlfma.cpp:50-695,
// integrate.cpp:7-69); the layouts are this library's own:
//   * every per-sample table is stored sample-fastest ("tap-major"), so a
//     wavefront reading one tap for 64 consecutive samples issues one
//     coalesced load;
//   * interpolation taps are contiguous modulo the source sample count, so
//     only the first index is stored;
//   * near-field blocks are stored basis-major [slot][basis][test];
//   * far-field source lists are not stored at all: the kernels enumerate the
//     <=27 interaction-list boxes from Morton coordinates.
#pragma once

#include <complex>
#include <cstdint>
#include <string>
#include <vector>

namespace isg {

typedef std::complex<double> cd;

struct PlanConfig {
  int level = 4;        // quadtree levels counting the root (mom.cpp:6)
  int box = 8;          // pixels per leaf-box side (mom.cpp:5); the kernels need 8
  double res = 0.1;     // pixel size in wavelengths (mom.cpp:4)
  double scale = 1.0;   // frequency scale, k0 = 2 pi scale (mom.cpp:136)
  int beta = 5;         // excess-bandwidth parameter (setup_mlfma.cpp:3)
  int ninter = 30;      // Lagrange taps (setup_mlfma.cpp:4)
  int numrx = 16;       // far-field receivers (mom.cpp:10)
  // 0: plane-wave transmitters and far-field receivers (the reference);
  // > 0: line-source transmitters and point receivers on a circle of this
  // radius (wavelengths), equispaced from angle 0
  double radius = 0.0;
  // false: the near blocks, P2M/L2P patterns, shifts and translation
  // operators are left to the device set-up (their host vectors stay empty)
  bool host_tables = true;
};

struct Plan {
  PlanConfig cfg;
  int L = 0, side = 0, N = 0, B = 0;  // levels, pixels per side, unknowns, leaf boxes
  double k0 = 0;
  std::vector<int> nclus, nsamp, nterm;
  std::vector<double> csize;

  // geometry
  std::vector<cd> pos;          // pixel centres in tree order
  std::vector<int> raster2tree; // raster index k*side+l -> tree index

  // tables (complex FP64; the FP32 path converts on upload)
  std::vector<cd> near;         // [9 slot][64 basis][64 test]
  std::vector<cd> coeffT;       // P2M [64 pixel][S_leaf]
  std::vector<cd> basisT;       // L2P [S_leaf][64 pixel]
  std::vector<std::vector<double>> interpT;  // level i (parent) [ninter][S_i]
  std::vector<std::vector<int>> interp0;     // first child index [S_i]
  std::vector<std::vector<double>> anterpT;  // level i (child) [2 ninter][S_i]
  std::vector<std::vector<int>> anterp0;     // first parent index [S_i]
  std::vector<std::vector<cd>> shiftM;       // level i [4][S_i]  child -> parent
  std::vector<std::vector<cd>> shiftL;       // level i [4][S_i]  parent -> child
  std::vector<std::vector<cd>> trans;        // level i>=2 [49][S_i]
  std::vector<double> rxT;      // [ninter][numrx]
  std::vector<int> rx0;         // [numrx]
  std::vector<double> rxaT;     // [2 ninter][S_0] receiver anterpolation weights
  std::vector<int> rxaI;        // [2 ninter][S_0] receiver indices (not contiguous)

  // counts for the work model
  long long near_pairs = 0;     // sum over leaf boxes of in-domain 3x3 neighbours
  std::vector<long long> far_pairs;  // per level

  std::string error;            // empty when the plan is valid
};

// Build the plan; returns false (with plan.error set) on invalid input.
bool build_plan(const PlanConfig& cfg, Plan& plan);

// Pixel-integrated Green's function (i/4)H0^(1) between two pixel centres:
// integrate.cpp:15-41.
cd green_pixel(double k0, double res, cd pt, cd pb);

// Morton helpers: cluster index <-> (column X, row Y) at a level, with the
// row bit the more significant of each pair (child = 4 par + 2 k + l,
// setup_mlfma.cpp:137).
inline uint32_t morton_spread(uint32_t v) {
  v &= 0x0000ffffu;
  v = (v | (v << 8)) & 0x00ff00ffu;
  v = (v | (v << 4)) & 0x0f0f0f0fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}
inline uint32_t morton_compact(uint32_t v) {
  v &= 0x55555555u;
  v = (v | (v >> 1)) & 0x33333333u;
  v = (v | (v >> 2)) & 0x0f0f0f0fu;
  v = (v | (v >> 4)) & 0x00ff00ffu;
  v = (v | (v >> 8)) & 0x0000ffffu;
  return v;
}
inline uint32_t morton_index(uint32_t X, uint32_t Y) { return (morton_spread(Y) << 1) | morton_spread(X); }

}  // namespace isg
