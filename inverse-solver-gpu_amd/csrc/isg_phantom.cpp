// isg_phantom.cpp -- seeded synthetic contrast profiles (raster order).
//
// The reference reads a real phantom from a text file (setup_born.cpp:103-113)
// that is not part of its tree.  These generators produce the seeded profiles
// the benchmark configurations name: Gaussian blobs and random ellipses.  The
// random stream is splitmix64 (Steele, Lea, Flood 2014), uniform doubles from
// the top 53 bits, so the same stream can be reproduced in any language.
#include <cmath>
#include <cstdint>

namespace {

struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double uniform(double a, double b) { return a + (b - a) * ((double)(next() >> 11) * 0x1.0p-53); }
};

}  // namespace

extern "C" {

// Sum of `n` Gaussian blobs: centres uniform in the inner 80% of the domain,
// sigma U(0.3, 1) wavelengths, amplitude U(0.1, 0.5) + i U(0.01, 0.05).
// Draw order per blob: cx, cy, sigma, amp_re, amp_im.  out: 2*side*side doubles.
int isg_phantom_blobs(int side, double res, int n, unsigned long long seed, double* out) {
  if (side < 1 || n < 0 || !out || !(res > 0)) return -1;
  double D = side * res;
  SplitMix64 rng{seed};
  for (long i = 0; i < 2L * side * side; i++) out[i] = 0;
  for (int b = 0; b < n; b++) {
    double cx = rng.uniform(-0.4 * D, 0.4 * D), cy = rng.uniform(-0.4 * D, 0.4 * D);
    double sg = rng.uniform(0.3, 1.0), ar = rng.uniform(0.1, 0.5), ai = rng.uniform(0.01, 0.05);
    for (int k = 0; k < side; k++)
      for (int l = 0; l < side; l++) {
        double x = -D / 2 + (l + 0.5) * res, y = D / 2 - (k + 0.5) * res;
        double e = std::exp(-((x - cx) * (x - cx) + (y - cy) * (y - cy)) / (2 * sg * sg));
        out[2 * ((long)k * side + l)] += ar * e;
        out[2 * ((long)k * side + l) + 1] += ai * e;
      }
  }
  return 0;
}

// `n` ellipses painted in order (later ones overwrite): centres uniform in the
// inner 80%, semi-axes U(0.02 D, 0.08 D), orientation U(0, pi), value
// U(vmin, vmax) + i imag.  Draw order per ellipse: cx, cy, a, b, theta, value.
int isg_phantom_ellipses(int side, double res, int n, unsigned long long seed, double vmin, double vmax, double imag,
                         double* out) {
  if (side < 1 || n < 0 || !out || !(res > 0)) return -1;
  double D = side * res;
  SplitMix64 rng{seed};
  for (long i = 0; i < 2L * side * side; i++) out[i] = 0;
  for (int e = 0; e < n; e++) {
    double cx = rng.uniform(-0.4 * D, 0.4 * D), cy = rng.uniform(-0.4 * D, 0.4 * D);
    double a = rng.uniform(0.02 * D, 0.08 * D), b = rng.uniform(0.02 * D, 0.08 * D);
    double th = rng.uniform(0.0, M_PI), v = rng.uniform(vmin, vmax);
    double c = std::cos(th), s = std::sin(th);
    for (int k = 0; k < side; k++)
      for (int l = 0; l < side; l++) {
        double x = -D / 2 + (l + 0.5) * res - cx, y = D / 2 - (k + 0.5) * res - cy;
        double u = (x * c + y * s) / a, w = (-x * s + y * c) / b;
        if (u * u + w * w <= 1.0) {
          out[2 * ((long)k * side + l)] = v;
          out[2 * ((long)k * side + l) + 1] = imag;
        }
      }
  }
  return 0;
}

}  // extern "C"
