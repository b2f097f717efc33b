// isg_device.hpp -- complex arithmetic and reduction helpers for gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace isg {

// Interleaved complex value; alignment lets a wavefront move one complex per
// lane with a single 8- or 16-byte load.
template <typename T>
struct alignas(2 * sizeof(T)) cplx {
  T x, y;
};

template <typename T>
__host__ __device__ __forceinline__ cplx<T> mkc(T a, T b) {
  cplx<T> r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) {
  return mkc<T>(a.x + b.x, a.y + b.y);
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) {
  return mkc<T>(a.x - b.x, a.y - b.y);
}
template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return mkc<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// acc + a*b as two fused multiply-adds per component (4 instructions per
// complex multiply-add; a separate product and sum would take 6)
template <typename T>
__device__ __forceinline__ cplx<T> cfma(cplx<T> a, cplx<T> b, cplx<T> acc) {
  return mkc<T>(fma(a.x, b.x, fma(-a.y, b.y, acc.x)), fma(a.x, b.y, fma(a.y, b.x, acc.y)));
}
template <typename T>
__device__ __forceinline__ cplx<T> cconj(cplx<T> a) {
  return mkc<T>(a.x, -a.y);
}
// conj(a) * b
template <typename T>
__device__ __forceinline__ cplx<T> cconjmul(cplx<T> a, cplx<T> b) {
  return mkc<T>(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
template <typename T>
__device__ __forceinline__ cplx<T> cscale(T s, cplx<T> a) {
  return mkc<T>(s * a.x, s * a.y);
}
template <typename To, typename Ti>
__device__ __forceinline__ cplx<To> ccast(cplx<Ti> a) {
  return mkc<To>((To)a.x, (To)a.y);
}

typedef cplx<double> zd;
typedef cplx<float> zf;

__host__ __device__ __forceinline__ zd zdiv(zd a, zd b) {
  double den = b.x * b.x + b.y * b.y;
  return mkc<double>((a.x * b.x + a.y * b.y) / den, (a.y * b.x - a.x * b.y) / den);
}
__host__ __device__ __forceinline__ double zabs(zd a) {
  return sqrt(a.x * a.x + a.y * a.y);
}

// Morton code of a (column X, row Y) pair, row bit more significant.
__host__ __device__ __forceinline__ unsigned spread2(unsigned v) {
  v &= 0x0000ffffu;
  v = (v | (v << 8)) & 0x00ff00ffu;
  v = (v | (v << 4)) & 0x0f0f0f0fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}
__host__ __device__ __forceinline__ unsigned compact2(unsigned v) {
  v &= 0x55555555u;
  v = (v | (v >> 1)) & 0x33333333u;
  v = (v | (v >> 2)) & 0x0f0f0f0fu;
  v = (v | (v >> 4)) & 0x00ff00ffu;
  v = (v | (v >> 8)) & 0x0000ffffu;
  return v;
}
__host__ __device__ __forceinline__ unsigned morton2(unsigned X, unsigned Y) {
  return (spread2(Y) << 1) | spread2(X);
}

// Sum of a double over the 64 lanes of a wavefront; lane 0 holds the result.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// A BiCGSTAB step done in the near-field kernel's epilogue instead of a
// separate pass over the operator output (vectors of storage type V):
//   mode 1: out = v = post(p, G y) and the partials of <r~, v>
//   mode 2: out = t = post(s, G y) and the partials of <t, s>, <t, t>
//   mode 3 (FP64 only): out = G y itself and the partials of |b|^2 and
//           |b - post(x, G y)|^2 -- the true-residual check of an iterate x;
//           the operator's input is then x itself, and P2M and the near field
//           form y = prep(x) while loading it
// post() is the contrast product of the solver (forward x - Gy, adjoint
// x - conj(g Gy)); partials go to part[rhs][slot][workgroup].
template <typename V>
struct NearPost {
  int mode = 0;
  const cplx<V>* x = nullptr;   // p (mode 1), s (mode 2) or the iterate (mode 3)
  const cplx<V>* rt = nullptr;  // r~ (mode 1)
  const cplx<V>* b = nullptr;   // right-hand side (mode 3)
  const zd* g = nullptr;
  long gs = 0;
  int herm = 0;
  double2* part = nullptr;
};

}  // namespace isg
