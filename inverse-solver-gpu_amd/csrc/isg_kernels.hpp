// isg_kernels.hpp -- HIP kernels of the MLFMA operator and the BiCGSTAB
// vector algebra, written for gfx950 (64-lane wavefronts, 256-thread blocks).
//
// Operator stages (reference counterparts in parentheses):
//   k_p2m_mfma   leaf radiation patterns on the matrix cores (kernels.cu:6-30)
//   k_m2m_lds    child -> parent interpolate+shift  (kernels.cu:83-136)
//   k_m2l_tile   diagonal translation               (kernels.cu:239-266)
//   k_l2l_lds    parent -> child shift+anterpolate  (kernels.cu:137-238)
//   k_near_mfma  near field + leaf receiving patterns on the matrix cores
//                                                (kernels.cu:32-81, :292-319)
//   k_rx_*    receiver interpolation / anterpolation (matvec_gpu.cu:351-361, :395-401)
// Grids are (work items / 256, right-hand sides); blockIdx.y selects the
// right-hand side and an optional `act` mask skips converged ones.
#pragma once

#include "isg_device.hpp"

namespace isg {

constexpr int kBlock = 256;

__device__ __forceinline__ bool rhs_off(const int* act, int rhs) { return act != nullptr && act[rhs] == 0; }

// XCD-aware workgroup order.  Workgroups are dealt to the 8 XCDs (each with
// its own L2) round-robin by linear id, so consecutive blockIdx.x run on
// different L2s.  When gridDim.x is a multiple of 8 (then x mod 8 is the
// XCD for every blockIdx.y), this bijection gives each XCD a contiguous
// range of logical ids: neighbouring tiles and the sample chunks of one tile,
// which read the same cache lines, then share one L2.
__device__ __forceinline__ int xcd_block() {
  const int x = blockIdx.x, n = gridDim.x;
  return (n & 7) ? x : (x & 7) * (n >> 3) + (x >> 3);
}

// Strided staging of n elements (element e = tid, tid + kBlock, ...): the
// loads of U consecutive strides, ld(e), are issued before any of them is
// stored with st(e, v), so each thread has U memory requests in flight
// instead of waiting for every load in turn before the next one issues.
// ld should only load (arithmetic on the loaded values belongs in st, or the
// compiler waits for each load before issuing the next).
template <typename T>
struct CPair {
  cplx<T> a, b;
};

template <int U, typename V, typename Ld, typename St>
__device__ __forceinline__ void stage(int n, Ld ld, St st) {
  for (int e0 = threadIdx.x; e0 < n; e0 += U * kBlock) {
    V v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int e = e0 + u * kBlock;
      if (e < n) v[u] = ld(e);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int e = e0 + u * kBlock;
      if (e < n) st(e, v[u]);
    }
  }
}

// Direct-from-cache interpolation kernels, one thread per target sample; used
// for root levels whose LDS staging would exceed 64 KB.
// parent[p][k] = sum_c shift[c][k] * sum_t w[t][k] child[4p+c][(i0[k]+t) mod Sc]
template <typename T>
__global__ void __launch_bounds__(kBlock) k_m2m(const cplx<T>* __restrict__ child, long cs, cplx<T>* __restrict__ par,
                                                long ps, const T* __restrict__ wT, const int* __restrict__ i0,
                                                const cplx<T>* __restrict__ shift, int Sp, int Sc, int np, int ntap,
                                                const int* __restrict__ act) {
  int rhs = blockIdx.y;
  if (rhs_off(act, rhs)) return;
  long idx = (long)blockIdx.x * kBlock + threadIdx.x;
  if (idx >= (long)np * Sp) return;
  int p = (int)(idx / Sp), k = (int)(idx - (long)p * Sp);
  int j0 = i0[k];
  cplx<T> acc = mkc<T>(0, 0);
  for (int c = 0; c < 4; c++) {
    const cplx<T>* ch = child + rhs * cs + (long)(4 * p + c) * Sc;
    cplx<T> red = mkc<T>(0, 0);
    int j = j0;
    for (int t = 0; t < ntap; t++) {
      T w = wT[t * Sp + k];
      cplx<T> v = ch[j];
      red.x += w * v.x;
      red.y += w * v.y;
      j = (j + 1 == Sc) ? 0 : j + 1;
    }
    acc = cfma(red, shift[c * Sp + k], acc);
  }
  par[rhs * ps + idx] = acc;
}

// child[c][k] (+)= sum_t a[t][k] * shiftL[c%4][j] * par[c/4][j], j = (j0[k]+t) mod Sp
template <typename T>
__global__ void __launch_bounds__(kBlock) k_l2l(const cplx<T>* __restrict__ par, long ps, cplx<T>* __restrict__ child,
                                                long cs, const T* __restrict__ aT, const int* __restrict__ j0s,
                                                const cplx<T>* __restrict__ shift, int Sp, int Sc, int np, int ntap,
                                                int accumulate, const int* __restrict__ act) {
  int rhs = blockIdx.y;
  if (rhs_off(act, rhs)) return;
  long idx = (long)blockIdx.x * kBlock + threadIdx.x;
  if (idx >= (long)4 * np * Sc) return;
  int c = (int)(idx / Sc), k = (int)(idx - (long)c * Sc);
  int p = c >> 2, cm = c & 3;
  const cplx<T>* pv = par + rhs * ps + (long)p * Sp;
  const cplx<T>* sh = shift + cm * Sp;
  cplx<T> acc = mkc<T>(0, 0);
  int j = j0s[k];
  for (int t = 0; t < ntap; t++) {
    T w = aT[t * Sc + k];
    cplx<T> v = cmul(sh[j], pv[j]);
    acc.x += w * v.x;
    acc.y += w * v.y;
    j = (j + 1 == Sp) ? 0 : j + 1;
  }
  cplx<T>* dst = child + rhs * cs + idx;
  *dst = accumulate ? cadd(*dst, acc) : acc;
}

// ---------------------------------------------------------------------------
// Matrix-core (MFMA) near field.  For a tile of NB Morton-consecutive leaf
// boxes the output r[box][m] (64 rows m = test pixel, NB columns) is one
// complex GEMM over K = 9*64 (near blocks times the gathered neighbour
// pixels) + S (leaf receiving pattern times the leaf locals, i.e. L2P), so
// the near field and L2P are fused and r is written once.
// Wave w owns rows 16w..16w+15; each 16x16 output block accumulates four real
// products (re*re, -im*im, re*im, im*re) with v_mfma_*_16x16x4.  Operand
// layout of the 16x16x4 instructions (checked on gfx950): A[i][k] in lane
// i + 16k, B[k][j] in lane j + 16k; the 16x16 result D[i][j] is in lane
// j + 16(i/4), register i%4 for f32 and in lane j + 16(i%4), register i/4
// for f64 -- Mfma<T>::row gives the row of (lane group g, register q).
// Every kernel puts the data vectors (boxes, clusters) in the A operand and
// the operator table in the B operand, so the 16 lanes of a group hold 16
// consecutive samples / pixels of the result and stores coalesce.
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static constexpr bool kRowsAcrossLanes = true;  // rows g + 4q: siblings across lane groups
  static __device__ __forceinline__ v4 mac(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int q) { return g + 4 * q; }
  // cr + i ci += d * a   (d data, A operand; a table, B operand; nai = -a.y)
  static __device__ __forceinline__ void cmac(cplx<double> d, cplx<double> a, double nai, v4& cr, v4& ci) {
    cr = mac(d.x, a.x, cr);
    cr = mac(d.y, nai, cr);
    ci = mac(d.x, a.y, ci);
    ci = mac(d.y, a.x, ci);
  }
  // three-product form (Gauss): p1 += dr ar, p2 += di ai, p3 += (dr + di)(ar + ai)
  // with asum = ar + ai; then d a = (p1 - p2) + i (p3 - p1 - p2)
  static __device__ __forceinline__ void cmac3(cplx<double> d, cplx<double> a, double asum, v4& p1, v4& p2, v4& p3) {
    p1 = mac(d.x, a.x, p1);
    p2 = mac(d.y, a.y, p2);
    p3 = mac(d.x + d.y, asum, p3);
  }
};
template <>
struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static constexpr bool kRowsAcrossLanes = false;  // rows 4g + q: siblings in one lane
  static __device__ __forceinline__ v4 mac(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int q) { return 4 * g + q; }
  static __device__ __forceinline__ void cmac(cplx<float> d, cplx<float> a, float nai, v4& cr, v4& ci) {
    cr = mac(d.x, a.x, cr);
    cr = mac(d.y, nai, cr);
    ci = mac(d.x, a.y, ci);
    ci = mac(d.y, a.x, ci);
  }
  static __device__ __forceinline__ void cmac3(cplx<float> d, cplx<float> a, float asum, v4& p1, v4& p2, v4& p3) {
    p1 = mac(d.x, a.x, p1);
    p2 = mac(d.y, a.y, p2);
    p3 = mac(d.x + d.y, asum, p3);
  }
};

// ---- BiCGSTAB helpers shared by the vector kernels and the near-field epilogue
constexpr int kSlots = 3;

// The contrast weight g[i] enters prep only in the forward product and post
// only in the adjoint one; it is loaded only there (herm is uniform per launch).
__device__ __forceinline__ zd prep_val(zd x, const zd* g, long i, bool herm) { return herm ? cconj(x) : cmul(x, g[i]); }
__device__ __forceinline__ zd post_val(zd x, zd gy, const zd* g, long i, bool herm) {
  return herm ? csub(x, cconj(cmul(g[i], gy))) : csub(x, gy);
}

// Block-level sum of up to 3 complex partials; thread 0 writes part[slot].
template <int NS>
__device__ __forceinline__ void block_store(zd (&v)[NS], double2* part, int nblk, int blk) {
  __shared__ double sh[NS * 2][kBlock / 64];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int s = 0; s < NS; s++) {
    double a = wave_sum(v[s].x), b = wave_sum(v[s].y);
    if (lane == 0) {
      sh[2 * s][w] = a;
      sh[2 * s + 1][w] = b;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int s = 0; s < NS; s++) {
      double a = 0, b = 0;
      for (int i = 0; i < kBlock / 64; i++) {
        a += sh[2 * s][i];
        b += sh[2 * s + 1][i];
      }
      part[s * nblk + blk] = make_double2(a, b);
    }
  }
}

#define ISG_VEC_LOOP(N) for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < (N); i += (long)gridDim.x * kBlock)

// Vector kernels are templated on the storage type V of the work vectors
// (double, or float for the mixed-precision solver); arithmetic and the dot
// product partials are FP64 either way, the solution x and the right-hand
// side b are always FP64.
template <typename V>
__device__ __forceinline__ zd ldv(const cplx<V>* a, long i) { return ccast<double>(a[i]); }
template <typename V>
__device__ __forceinline__ zd stv(cplx<V>* a, long i, zd v) {
  cplx<V> w = ccast<V>(v);
  a[i] = w;
  return ccast<double>(w);  // the value as stored
}

// LDS row length (complex elements) of the near/L2P tile: 64 pixels or S
// locals, plus one element of padding so 16 lanes reading 16 different rows
// hit 16 different banks.
// LDS elements of k_near_mfma<NB>: the (TX+2) x (TY+2) halo boxes of 65
// elements, later reused for NB rows of S + 1 leaf locals.
__host__ __device__ inline int near_lds_elems(int NB, int S) {
  int tx = (NB == 16) ? 4 : 8, ty = NB / tx;
  int halo = (tx + 2) * (ty + 2) * 65, l2p = NB * (S + 1);
  return halo > l2p ? halo : l2p;
}

template <typename T, typename TI, int NB, int POST>
__global__ void __launch_bounds__(kBlock) k_near_mfma(const cplx<TI>* __restrict__ y, long ys,
                                                      const cplx<T>* __restrict__ loc, long ls,
                                                      cplx<TI>* __restrict__ r, long rs,
                                                      const cplx<T>* __restrict__ near,
                                                      const cplx<T>* __restrict__ basisT, int S, int nbs,
                                                      int with_near, const int* __restrict__ act,
                                                      NearPost<TI> post) {
  typedef typename Mfma<T>::v4 v4;
  constexpr int NN = NB / 16;
  // Morton-consecutive NB boxes form a TX x TY block; with its one-box halo
  // the nine near slots of every box in the tile lie in HX x HY boxes, staged
  // once (each input pixel enters LDS once per tile instead of once per slot).
  constexpr int TX = (NB == 16) ? 4 : 8, TY = NB / TX, HX = TX + 2, HY = TY + 2;
  constexpr int HROW = 65;  // 64 pixels + 1 padding element
  constexpr int PRE = sizeof(T) == 4 ? 8 : 4;  // L2P table values loaded ahead per MFMA group
  extern __shared__ char smem_raw[];
  cplx<T>* tile = reinterpret_cast<cplx<T>*>(smem_raw);
  const int row = S + 1;    // L2P rows: S locals + 1 padding element
  int rhs = blockIdx.y;
  if (rhs_off(act, rhs)) return;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const long b0 = (long)xcd_block() * NB;
  // complex products in the three-product form (Mfma::cmac3)
  v4 p1[NN], p2[NN], p3[NN];
#pragma unroll
  for (int j = 0; j < NN; j++) {
    p1[j] = v4{0, 0, 0, 0};
    p2[j] = v4{0, 0, 0, 0};
    p3[j] = v4{0, 0, 0, 0};
  }
  const cplx<TI>* yr = y + rhs * ys;
  if (with_near) {
    const int X0 = (int)compact2((unsigned)b0), Y0 = (int)compact2((unsigned)b0 >> 1);
    // pixel offset of each halo box (Morton index * 64), -1 outside the grid:
    // one Morton conversion per box instead of one per pixel
    __shared__ int hoff[HX * HY];
    for (int h = tid; h < HX * HY; h += kBlock) {
      const int sx = X0 - 1 + h % HX, sy = Y0 - 1 + h / HX;
      hoff[h] = (sx >= 0 && sy >= 0 && sx < nbs && sy < nbs) ? (int)morton2(sx, sy) * 64 : -1;
    }
    __syncthreads();
    stage<sizeof(T) == 4 ? 8 : 4, cplx<TI>>(
        HX * HY * 64,
        [&](int idx) {
          const int off = hoff[idx >> 6];
          cplx<TI> v = mkc<TI>(0, 0);
          if (off >= 0) {
            v = yr[(long)off + (idx & 63)];
            // residual check: the input is the iterate x, contrast-weighted here
            if constexpr (POST == 3) v = prep_val(v, post.g, rhs * post.gs + off + (idx & 63), post.herm != 0);
          }
          return v;
        },
        [&](int idx, cplx<TI> v) { tile[(idx >> 6) * HROW + (idx & 63)] = ccast<T>(v); });
    // halo row of column 16 j + li for the centre slot
    int hc[NN];
#pragma unroll
    for (int j = 0; j < NN; j++) {
      unsigned c = (unsigned)(16 * j + li);
      hc[j] = ((int)compact2(c >> 1) + 1) * HX + (int)compact2(c) + 1;
    }
    __syncthreads();
    for (int s = 0; s < 9; s++) {
      const int off = (s / 3 - 1) * HX + (s % 3 - 1);
      const cplx<T>* A = near + (long)s * 4096 + 16 * w + li;
#pragma unroll 8
      for (int kk = 0; kk < 16; kk++) {
        const int k = 4 * kk + lk;
        cplx<T> a = A[k * 64];
        T asum = a.x + a.y;
#pragma unroll
        for (int j = 0; j < NN; j++) {
          cplx<T> bv = tile[(hc[j] + off) * HROW + k];
          Mfma<T>::cmac3(bv, a, asum, p1[j], p2[j], p3[j]);
        }
      }
    }
  }
  // leaf locals -> pixels (L2P) as the last K block
  __syncthreads();
  const cplx<T>* lr = loc + rhs * ls + b0 * S;
  {
    // element idx = j * S + k goes to tile[j * (S + 1) + k] = tile[idx + j];
    // j is stepped rather than divided out of idx
    constexpr int U = sizeof(T) == 4 ? 8 : 4;
    const int n = NB * S, dj = kBlock / S, dk = kBlock - dj * S;
    int jb = tid / S, kq = tid - jb * S;
    for (int e0 = tid; e0 < n; e0 += U * kBlock) {
      cplx<T> v[U];
      int ju[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int e = e0 + u * kBlock;
        ju[u] = jb;
        v[u] = lr[e < n ? e : n - 1];  // unconditional (clamped) load: the array stays in registers
        kq += dk;
        jb += dj;
        if (kq >= S) {
          kq -= S;
          jb++;
        }
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int e = e0 + u * kBlock;
        if (e < n) tile[e + ju[u]] = v[u];
      }
    }
  }
  __syncthreads();
  const cplx<T>* A = basisT + 16 * w + li;
  const int nk = S / 4;  // S is a multiple of 4
#pragma unroll 1
  for (int k0 = 0; k0 < nk; k0 += PRE) {
    cplx<T> a[PRE];
#pragma unroll
    for (int u = 0; u < PRE; u++)
      if (k0 + u < nk) a[u] = A[(4 * (k0 + u) + lk) * 64];
#pragma unroll
    for (int u = 0; u < PRE; u++) {
      if (k0 + u < nk) {
        const int k = 4 * (k0 + u) + lk;
        const T asum = a[u].x + a[u].y;
#pragma unroll
        for (int j = 0; j < NN; j++) {
          cplx<T> bv = tile[(16 * j + li) * row + k];
          Mfma<T>::cmac3(bv, a[u], asum, p1[j], p2[j], p3[j]);
        }
      }
    }
  }
  // D[box][m]: lane li holds pixel m = 16w + li.  The product goes through
  // LDS (the L2P tile is free once every wave has passed the barrier) so
  // that each thread then takes consecutive elements of the workgroup's
  // contiguous output range [b0*64, (b0+NB)*64): coalesced stores, and for a
  // post step coalesced loads and few live registers.
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NN; j++) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const T re = p1[j][q] - p2[j][q], im = p3[j][q] - p1[j][q] - p2[j][q];
      tile[(16 * j + Mfma<T>::row(lk, q)) * 64 + 16 * w + li] = mkc<T>(re, im);
    }
  }
  __syncthreads();
  if constexpr (POST == 0) {
    cplx<TI>* out = r + rhs * rs + b0 * 64;
#pragma unroll 4
    for (int e = tid; e < NB * 64; e += kBlock) out[e] = ccast<TI>(tile[e]);
  } else {
  // BiCGSTAB post step (modes 1, 2) or residual check (mode 3, FP64) on the
  // operator output, rounded to the vector type TI first, as a separate pass
  // reading the stored product back would see it.
  const long o = rhs * rs + b0 * 64;
  const zd* gr = post.g + rhs * post.gs + b0 * 64;
  zd acc[2] = {mkc<double>(0, 0), mkc<double>(0, 0)};
  struct In {
    cplx<TI> x, aux;  // aux: r~ (mode 1) or b (mode 3)
    zd g;
  };
  const bool herm = post.herm != 0;
  stage<2, In>(
      NB * 64,
      [&](int e) {
        In v;
        v.x = post.x[o + e];
        if (POST == 1) v.aux = post.rt[o + e];
        if (POST == 3) v.aux = post.b[o + e];
        if (herm) v.g = gr[e];  // the forward post step does not use g
        return v;
      },
      [&](int e, In v) {
        const cplx<T> gt = tile[e];
        const zd gy = mkc<double>((double)(TI)gt.x, (double)(TI)gt.y);
        const zd xv = ccast<double>(v.x);
        const zd px = herm ? csub(xv, cconj(cmul(v.g, gy))) : csub(xv, gy);
        if constexpr (POST == 3) {
          r[o + e] = ccast<TI>(gt);  // the product itself, as the plain kernel stores it
          const zd bv = ccast<double>(v.aux), rv = csub(bv, px);
          acc[0].x += bv.x * bv.x + bv.y * bv.y;
          acc[1].x += rv.x * rv.x + rv.y * rv.y;
        } else {
          const zd res = stv(r, o + e, px);
          if (POST == 1) {
            const zd a = ccast<double>(v.aux);
            acc[0].x += a.x * res.x + a.y * res.y;
            acc[0].y += a.x * res.y - a.y * res.x;
          } else {
            acc[0].x += res.x * xv.x + res.y * xv.y;
            acc[0].y += res.x * xv.y - res.y * xv.x;
            acc[1].x += res.x * res.x + res.y * res.y;
          }
        }
      });
    block_store<2>(acc, post.part + (long)rhs * kSlots * gridDim.x, gridDim.x, blockIdx.x);
  }
}

// ---------------------------------------------------------------------------
// Sample-chunked LDS interpolation (upper levels: few clusters, many samples).
// A workgroup owns one parent (or G small parents) and a chunk of <= 256
// target samples.  Taps of consecutive targets advance monotonically (mod the
// source count), so the chunk needs one contiguous window of each source
// vector: window start = first[k0], offset of target k = (first[k] -
// first[k0]) mod Ssrc.  All four windows (children, or shifted parent copies)
// go to LDS once; each thread then forms one target from LDS.

// M2M: parent[p][k] = sum_c shiftM[c][k] sum_t w[t][k] child[4p+c][first[k]+t]
template <typename T>
__global__ void __launch_bounds__(kBlock) k_m2m_chunk(const cplx<T>* __restrict__ child, long cs,
                                                      cplx<T>* __restrict__ par, long ps, const T* __restrict__ wT,
                                                      const int* __restrict__ i0, const cplx<T>* __restrict__ shift,
                                                      int Sp, int Sc, int G, int kch, int nch, int ntap,
                                                      const int* __restrict__ act) {
  extern __shared__ char smem_raw[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
  const int rhs = blockIdx.y;
  if (rhs_off(act, rhs)) return;
  const int pg = blockIdx.x / nch, ch = blockIdx.x - pg * nch;
  const long p0 = (long)pg * G;
  const int k0 = ch * kch, kn = min(kch, Sp - k0);
  const int base = i0[k0];
  int W = i0[k0 + kn - 1] - base;
  if (W < 0) W += Sc;
  W += ntap;
  const cplx<T>* cr = child + rhs * cs + 4 * p0 * Sc;
  stage<sizeof(T) == 4 ? 8 : 4, cplx<T>>(
      4 * G * W,
      [&](int e) {
        int v = e / W, t = e - v * W;  // v = 4 g + c (consecutive children)
        int j = base + t;
        while (j >= Sc) j -= Sc;
        return cr[(long)v * Sc + j];
      },
      [&](int e, cplx<T> v) { lds[e] = v; });
  __syncthreads();
  int o = threadIdx.x;
  if (o >= G * kn) return;
  int g = o / kn, k = k0 + (o - g * kn);
  int d = i0[k] - base;
  if (d < 0) d += Sc;
  cplx<T> acc = mkc<T>(0, 0);
  for (int c = 0; c < 4; c++) {
    const cplx<T>* src = lds + (4 * g + c) * W + d;
    cplx<T> red = mkc<T>(0, 0);
    for (int t = 0; t < ntap; t++) {
      T w = wT[t * Sp + k];
      red.x += w * src[t].x;
      red.y += w * src[t].y;
    }
    acc = cfma(red, shift[c * Sp + k], acc);
  }
  par[rhs * ps + (p0 + g) * Sp + k] = acc;
}

// L2L: child[4p+c][k] (+)= sum_t a[t][k] (shiftL[c] o par[p])[first[k]+t]
template <typename T>
__global__ void __launch_bounds__(kBlock) k_l2l_chunk(const cplx<T>* __restrict__ par, long ps,
                                                      cplx<T>* __restrict__ child, long cs, const T* __restrict__ aT,
                                                      const int* __restrict__ j0s, const cplx<T>* __restrict__ shift,
                                                      int Sp, int Sc, int kch, int nch, int ntap, int accumulate,
                                                      const int* __restrict__ act) {
  extern __shared__ char smem_raw[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_raw);
  const int rhs = blockIdx.y;
  if (rhs_off(act, rhs)) return;
  const int p = blockIdx.x / nch, ch = blockIdx.x - p * nch;
  const int k0 = ch * kch, kn = min(kch, Sc - k0);
  const int base = j0s[k0];
  int W = j0s[k0 + kn - 1] - base;
  if (W < 0) W += Sp;
  W += ntap;
  const cplx<T>* pv = par + rhs * ps + (long)p * Sp;
  stage<4, CPair<T>>(
      4 * W,
      [&](int e) {
        int c = e / W, t = e - c * W;
        int j = base + t;
        while (j >= Sp) j -= Sp;
        return CPair<T>{shift[c * Sp + j], pv[j]};
      },
      [&](int e, CPair<T> v) { lds[e] = cmul(v.a, v.b); });
  __syncthreads();
  for (int o = threadIdx.x; o < 4 * kn; o += kBlock) {
    int c = o / kn, k = k0 + (o - c * kn);
    int d = j0s[k] - base;
    if (d < 0) d += Sp;
    const cplx<T>* src = lds + c * W + d;
    cplx<T> red = mkc<T>(0, 0);
    for (int t = 0; t < ntap; t++) {
      T w = aT[t * Sc + k];
      red.x += w * src[t].x;
      red.y += w * src[t].y;
    }
    cplx<T>* dst = child + rhs * cs + (4L * p + c) * Sc + k;
    *dst = accumulate ? cadd(*dst, red) : red;
  }
}

// Interpolation (M2M) and anterpolation (L2L) on the matrix cores, for the
// levels with many clusters.  Out[k][col] = sum_j A[k][j] In[j][col], A the
// real interpolator (30 taps per row, M2M) or anterpolator (60 taps, L2L),
// columns = the child vectors (4 per parent) of NC/4 consecutive parents.
// A is banded: the taps of 16 consecutive rows k all fall in one window of
// KB <= K inputs starting at s0[rb] (modulo K), so each 16-row block is a
// 16 x KB dense block ("band block"), stored K-major as AB[rb][kk][16].
//   M2M (MODE 0): In = children (contiguous); the epilogue multiplies by
//                 shiftM[c][k] and sums the 4 sibling columns, then writes
//                 parent[p][k].
//   L2L (MODE 1): In = shiftL[c] o parent, staged per column; Out is added to
//                 (or assigned to) child[4p+c][k].
// The KB/4 table fragments of a block are loaded into registers before its
// MFMA chain; the data columns come from LDS (one padding element per column).
template <typename T, int MODE, int NCB, int KQ>
__global__ void __launch_bounds__(kBlock) k_interp_band(const cplx<T>* __restrict__ src, long ss,
                                                        cplx<T>* __restrict__ dst, long ds, const T* __restrict__ AB,
                                                        const int* __restrict__ s0v, int KB,
                                                        const cplx<T>* __restrict__ shift, int Mrows, int K,
                                                        int accumulate, const int* __restrict__ act) {
  typedef typename Mfma<T>::v4 v4;
  constexpr int NC = 16 * NCB;
  extern __shared__ char smem_raw[];
  cplx<T>* tile = reinterpret_cast<cplx<T>*>(smem_raw);
  const int rhs = blockIdx.y;
  if (rhs_off(act, rhs)) return;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int row = K + 1;
  const long p0 = (long)xcd_block() * (NC / 4);
  // M2M: element e = col * K + j of the NC x K block of children goes to
  // tile[col * (K + 1) + j] = tile[e + col].  L2L: element e = p * K + j of the
  // NC/4 parents is loaded once and gives the 4 children's columns
  // tile[(4 p + c) * (K + 1) + j] = shiftL[c][j] parent[p][j].  Each thread
  // starts at e = tid and advances by kBlock; (col, j) are stepped rather than
  // divided out of e, and the loads of U strides are issued before the first
  // store.
  if constexpr (MODE == 0) {
    constexpr int U = sizeof(T) == 4 ? 8 : 4;
    const int n = NC * K, dcol = kBlock / K, dj = kBlock - dcol * K;
    int col = tid / K, j = tid - col * K;
    const cplx<T>* s = src + rhs * ss + 4 * p0 * K;
    for (int e0 = tid; e0 < n; e0 += U * kBlock) {
      cplx<T> a[U];
      int cu[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int e = e0 + u * kBlock;
        cu[u] = col;
        a[u] = s[e < n ? e : n - 1];  // NC consecutive children of length K (clamped load)
        j += dj;
        col += dcol;
        if (j >= K) {
          j -= K;
          col++;
        }
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int e = e0 + u * kBlock;
        if (e < n) tile[e + cu[u]] = a[u];
      }
    }
  } else {
    constexpr int U = 4;
    const int n = (NC / 4) * K, dp = kBlock / K, dj = kBlock - dp * K;
    int pp = tid / K, j = tid - pp * K;
    const cplx<T>* s = src + rhs * ss + p0 * K;
    for (int e0 = tid; e0 < n; e0 += U * kBlock) {
      cplx<T> a[U];
      int pu[U], ju[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int e = e0 + u * kBlock;
        pu[u] = pp;
        ju[u] = j;
        a[u] = s[e < n ? e : n - 1];  // NC/4 consecutive parents of length K (clamped load)
        j += dj;
        pp += dp;
        if (j >= K) {
          j -= K;
          pp++;
        }
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int e = e0 + u * kBlock;
        if (e < n) {
          cplx<T>* t = tile + 4 * pu[u] * row + ju[u];
#pragma unroll
          for (int c = 0; c < 4; c++) t[c * row] = cmul(shift[c * K + ju[u]], a[u]);
        }
      }
    }
  }
  __syncthreads();
  const int nrb = (Mrows + 15) / 16;
  const int nkq = KB >> 2;
  // (row block, column block) items spread evenly over the 4 waves; the M2M
  // with two column blocks takes both of a row block in one item, so that its
  // band rows and sibling shifts (which do not depend on the column) are
  // loaded once for both
  constexpr bool RB_ITEMS = MODE == 0 && NCB == 2;
  constexpr int CPI = RB_ITEMS ? NCB : 1;  // column blocks per item
  for (int item = w; item < nrb * NCB / CPI; item += 4) {
    const int rb = RB_ITEMS ? item : item / NCB, cb0 = RB_ITEMS ? 0 : item - rb * NCB;
    const T* ab = AB + (long)rb * KB * 16;
    T a[KQ];
#pragma unroll
    for (int q = 0; q < KQ; q++) a[q] = (q < nkq) ? ab[(4 * q + lk) * 16 + li] : (T)0;
    const int s0 = s0v[rb];
    // the epilogue's operands (M2M: sibling shifts; L2L: the child locals it
    // adds to) are loaded here, so their latency overlaps the MFMA chain
    const int k = 16 * rb + li;
    cplx<T> pre[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      pre[q] = mkc<T>(0, 0);
      if (k < Mrows) {
        if (MODE == 0) {
          if (!Mfma<T>::kRowsAcrossLanes || q == 0)
            pre[q] = shift[(Mfma<T>::kRowsAcrossLanes ? lk : q) * Mrows + k];
        } else if (accumulate) {
          pre[q] = dst[rhs * ds + (4 * p0 + 16 * cb0 + Mfma<T>::row(lk, q)) * (long)Mrows + k];
        }
      }
    }
#pragma unroll
    for (int cc = 0; cc < CPI; cc++) {
      const int cb = cb0 + cc;
      const cplx<T>* tcol = tile + (16 * cb + li) * row;
      v4 cr0 = v4{0, 0, 0, 0}, ci0 = v4{0, 0, 0, 0};
#pragma unroll
      for (int q = 0; q < KQ; q++) {
        if (q < nkq) {
          int j = s0 + 4 * q + lk;
          if (j >= K) j -= K;
          const cplx<T> bv = tcol[j];
          cr0 = Mfma<T>::mac(bv.x, a[q], cr0);
          ci0 = Mfma<T>::mac(bv.y, a[q], ci0);
        }
      }
      // D[col][k]: lane li holds target sample k = 16 rb + li; col = 16 cb + row(lk, q)
      if (MODE == 0) {
        // parent = sum over the 4 sibling columns of shiftM[c][k] * D[col][k]
        if (Mfma<T>::kRowsAcrossLanes) {
          // col = 16cb + lk + 4q: sibling c = lk, parent 4cb + q
          const cplx<T> s1 = pre[0];
#pragma unroll
          for (int q = 0; q < 4; q++) {
            T vr = cr0[q], vi = ci0[q];
            T xr = vr * s1.x - vi * s1.y, xi = vr * s1.y + vi * s1.x;
            xr += __shfl_xor(xr, 16, 64);
            xi += __shfl_xor(xi, 16, 64);
            xr += __shfl_xor(xr, 32, 64);
            xi += __shfl_xor(xi, 32, 64);
            if (lk == 0 && k < Mrows) dst[rhs * ds + (p0 + 4 * cb + q) * Mrows + k] = mkc<T>(xr, xi);
          }
        } else {
          // col = 16cb + 4lk + q: sibling c = q, parent 4cb + lk
          T xr = 0, xi = 0;
#pragma unroll
          for (int q = 0; q < 4; q++) {
            const cplx<T> sq = pre[q];
            T vr = cr0[q], vi = ci0[q];
            xr += vr * sq.x - vi * sq.y;
            xi += vr * sq.y + vi * sq.x;
          }
          if (k < Mrows) dst[rhs * ds + (p0 + 4 * cb + lk) * Mrows + k] = mkc<T>(xr, xi);
        }
      } else if (k < Mrows) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int col = 16 * cb + Mfma<T>::row(lk, q);
          cplx<T>* o = dst + rhs * ds + (4 * p0 + col) * (long)Mrows + k;
          cplx<T> v = mkc<T>(cr0[q], ci0[q]);
          *o = accumulate ? cadd(pre[q], v) : v;
        }
      }
    }
  }
}

// M2L staging for a TS x TS tile of target boxes at corner (X0, Y0) and
// samples k0 .. k0+KC-1: the multipoles of the (TS+4)^2 boxes around it
// (zero outside the level or past S) and the 49 translation operators.
template <typename T, int TS, int KC>
__device__ __forceinline__ void m2l_stage(cplx<T> (*src)[KC], cplx<T> (*tr)[KC], const cplx<T>* a,
                                          const cplx<T>* trans, int X0, int Y0, int k0, int nb, int S) {
  constexpr int SB = TS + 4, NS = SB * SB * KC;
  constexpr int US = NS / kBlock < 9 ? (NS + kBlock - 1) / kBlock : 9;
  // offset of each halo box's multipole (Morton index * S), -1 outside the
  // level: one Morton conversion per box instead of one per element
  __shared__ int boff[SB * SB];
  for (int b = threadIdx.x; b < SB * SB; b += kBlock) {
    const int sx = X0 - 2 + b % SB, sy = Y0 - 2 + b / SB;
    boff[b] = (sx >= 0 && sy >= 0 && sx < nb && sy < nb) ? (int)morton2(sx, sy) * S : -1;
  }
  __syncthreads();
  // kBlock is a multiple of KC: each thread keeps one sample kk
  const int kk = threadIdx.x % KC;
  const bool kin = k0 + kk < S;
  stage<US, cplx<T>>(
      NS,
      [&](int e) {
        const int off = boff[e / KC];
        cplx<T> v = mkc<T>(0, 0);
        if (off >= 0 && kin) v = a[(long)off + k0 + kk];
        return v;
      },
      [&](int e, cplx<T> v) { src[e / KC][kk] = v; });
  stage<(49 * KC + kBlock - 1) / kBlock, cplx<T>>(
      49 * KC, [&](int e) { return kin ? trans[(long)(e / KC) * S + k0 + kk] : mkc<T>(0, 0); },
      [&](int e, cplx<T> v) { tr[e / KC][kk] = v; });
}

// Tiled M2L.  A workgroup owns a TS x TS block of target boxes (Morton
// consecutive) and KC consecutive samples.  The interaction lists of the
// block lie in the (TS+4)^2 block of boxes around it; those multipoles and
// the 49 translation operators for the KC samples are staged in LDS once, so
// each multipole sample is read from memory (TS+4)^2/TS^2 times instead of
// once per interaction (<= 27).
template <typename T, int TS, int KC>
__global__ void __launch_bounds__(kBlock) k_m2l_tile(const cplx<T>* __restrict__ agm, cplx<T>* __restrict__ agl,
                                                     long st, const cplx<T>* __restrict__ trans, int lev, int S,
                                                     const int* __restrict__ act) {
  constexpr int SB = TS + 4;
  __shared__ cplx<T> src[SB * SB][KC];
  __shared__ cplx<T> tr[49][KC];
  const int rhs = blockIdx.y;
  if (rhs_off(act, rhs)) return;
  const int nb = 1 << lev;
  const int nchunk = (S + KC - 1) / KC;
  const int lb = xcd_block();
  const int tile = lb / nchunk, chunk = lb - tile * nchunk;
  const int k0 = chunk * KC;
  const unsigned m0 = (unsigned)tile * TS * TS;
  const int X0 = (int)compact2(m0), Y0 = (int)compact2(m0 >> 1);
  m2l_stage<T, TS, KC>(src, tr, agm + rhs * st, trans, X0, Y0, k0, nb, S);
  __syncthreads();
  for (int o = threadIdx.x; o < TS * TS * KC; o += kBlock) {
    int tl = o / KC, kk = o - tl * KC;
    if (k0 + kk >= S) continue;
    unsigned m = m0 + tl;
    int X = (int)compact2(m), Y = (int)compact2(m >> 1);
    int bx = (X >> 1) * 2 - 2, by = (Y >> 1) * 2 - 2;
    cplx<T> acc = mkc<T>(0, 0);
    for (int sy = by; sy < by + 6; sy++) {
      if (sy < 0 || sy >= nb) continue;
      int dy = sy - Y;
      for (int sx = bx; sx < bx + 6; sx++) {
        if (sx < 0 || sx >= nb) continue;
        int dx = sx - X;
        if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1) continue;
        acc = cfma(tr[7 * (dy + 3) + dx + 3][kk], src[(sy - Y0 + 2) * SB + (sx - X0 + 2)][kk], acc);
      }
    }
    agl[rhs * st + (long)m * S + k0 + kk] = acc;
  }
}

// M2L with the translation operators in registers.  Staging as k_m2l_tile
// (8x8 target boxes, their 12x12 source halo, KC samples).  A target box's
// far list depends only on its position class (X & 1, Y & 1): the 27 boxes
// of the parent's 6x6 neighbourhood outside its own 3x3, at offsets fixed
// per class.  Each thread owns one sample kk and one class, loads that
// class's 27 operators once and sweeps 64 / (256 / KC) targets of the class
// with compile-time LDS offsets; source boxes outside the domain are staged
// as zeros, so no bounds tests remain.
template <typename T, int PX, int PY, int KC, int NT>
__device__ __forceinline__ void m2l_class(const cplx<T> (*src)[KC], const cplx<T> (*tr)[KC], int kk, int part,
                                          cplx<T>* __restrict__ out, unsigned m0, int S, int k0) {
  constexpr int SB = 12;
  cplx<T> t[27];
  {
    int j = 0;
#pragma unroll
    for (int dy = -2 - PY; dy <= 3 - PY; dy++)
#pragma unroll
      for (int dx = -2 - PX; dx <= 3 - PX; dx++)
        if (dx < -1 || dx > 1 || dy < -1 || dy > 1) t[j++] = tr[7 * (dy + 3) + dx + 3][kk];
  }
  // the class's 16 targets (2i + PX, 2j + PY), i, j = 0..3, numbered 4j + i;
  // this thread takes numbers part * NT .. part * NT + NT - 1
#pragma unroll 1
  for (int q = 0; q < NT; q++) {
    const int tt = part * NT + q;
    const int i = tt & 3, jj = tt >> 2;
    const int lx = 2 * i + PX, ly = 2 * jj + PY;
    const int base = (ly + 2) * SB + lx + 2;
    cplx<T> acc = mkc<T>(0, 0);
    int j = 0;
#pragma unroll
    for (int dy = -2 - PY; dy <= 3 - PY; dy++)
#pragma unroll
      for (int dx = -2 - PX; dx <= 3 - PX; dx++)
        if (dx < -1 || dx > 1 || dy < -1 || dy > 1) acc = cfma(t[j++], src[base + dy * SB + dx][kk], acc);
    if (k0 + kk < S) out[(long)(m0 + morton2((unsigned)lx, (unsigned)ly)) * S + k0 + kk] = acc;
  }
}

template <typename T, int KC>
__global__ void __launch_bounds__(kBlock) k_m2l_reg(const cplx<T>* __restrict__ agm, cplx<T>* __restrict__ agl,
                                                    long st, const cplx<T>* __restrict__ trans, int lev, int S,
                                                    const int* __restrict__ act) {
  constexpr int TS = 8, SB = TS + 4;
  constexpr int G = kBlock / KC;      // thread groups: 4 classes x G/4 parts
  constexpr int NT = 16 / (G / 4);    // targets per thread (16 per class per tile)
  static_assert(G % 4 == 0 && NT >= 1 && 16 % (G / 4) == 0, "KC must divide the tile");
  __shared__ cplx<T> src[SB * SB][KC];
  __shared__ cplx<T> tr[49][KC];
  const int rhs = blockIdx.y;
  if (rhs_off(act, rhs)) return;
  const int nb = 1 << lev;
  const int nchunk = (S + KC - 1) / KC;
  const int lb = xcd_block();
  const int tile = lb / nchunk, chunk = lb - tile * nchunk;
  const int k0 = chunk * KC;
  const unsigned m0 = (unsigned)tile * TS * TS;
  const int X0 = (int)compact2(m0), Y0 = (int)compact2(m0 >> 1);
  m2l_stage<T, TS, KC>(src, tr, agm + rhs * st, trans, X0, Y0, k0, nb, S);
  __syncthreads();
  const int kk = threadIdx.x % KC, grp = threadIdx.x / KC;
  const int cls = grp & 3, part = grp >> 2;
  cplx<T>* out = agl + rhs * st;
  switch (cls) {
    case 0: m2l_class<T, 0, 0, KC, NT>(src, tr, kk, part, out, m0, S, k0); break;
    case 1: m2l_class<T, 1, 0, KC, NT>(src, tr, kk, part, out, m0, S, k0); break;
    case 2: m2l_class<T, 0, 1, KC, NT>(src, tr, kk, part, out, m0, S, k0); break;
    default: m2l_class<T, 1, 1, KC, NT>(src, tr, kk, part, out, m0, S, k0); break;
  }
}

// M2L per parent: a workgroup owns 4x4 parents (8x8 target boxes) and 16
// samples; a thread owns one parent and one sample.  The far lists of the
// four children all lie in the parent's 6x6 neighbourhood of child-level
// boxes: the thread loads those 36 sources into registers once, then walks
// the 49 translation offsets, reading each operator once and applying it to
// every child for which that offset is a far interaction (27 per child).
// LDS reads per thread: 36 + 49 for 108 complex multiply-adds.
constexpr int kParTS = 8, kParSB = kParTS + 4, kParKC = 16;

// Workgroup coordinates shared by k_m2l_par and k_m2l_leaf: 8x8 boxes
// starting at Morton index m0 (corner X0, Y0), samples k0 .. k0+15.
struct ParBlock {
  int chunk, k0, X0, Y0;
  unsigned m0;
  __device__ __forceinline__ ParBlock(int S) {
    const int nchunk = (S + kParKC - 1) / kParKC;
    const int lb = xcd_block();
    const int tile = lb / nchunk;
    chunk = lb - tile * nchunk;
    k0 = chunk * kParKC;
    m0 = (unsigned)tile * kParTS * kParTS;
    X0 = (int)compact2(m0);
    Y0 = (int)compact2(m0 >> 1);
  }
};

// the parent's 6x6 source neighbourhood (first child at tile position lx0, ly0)
template <typename T>
__device__ __forceinline__ void par_load(const cplx<T> (*src)[kParKC], int lx0, int ly0, int kk, cplx<T> (&sv)[6][6]) {
#pragma unroll
  for (int dy = 0; dy < 6; dy++)
#pragma unroll
    for (int dx = 0; dx < 6; dx++) sv[dy][dx] = src[(ly0 + dy) * kParSB + lx0 + dx][kk];
}

// acc[c] = sum over the 27 far offsets of child c of T_offset * source
template <typename T>
__device__ __forceinline__ void par_translate(const cplx<T> (*tr)[kParKC], int kk, const cplx<T> (&sv)[6][6],
                                              cplx<T> (&acc)[4]) {
#pragma unroll
  for (int c = 0; c < 4; c++) acc[c] = mkc<T>(0, 0);
#pragma unroll
  for (int ey = -3; ey <= 3; ey++)
#pragma unroll
    for (int ex = -3; ex <= 3; ex++) {
      const cplx<T> t = tr[7 * (ey + 3) + ex + 3][kk];
      if (ex >= -1 && ex <= 1 && ey >= -1 && ey <= 1) continue;  // near offsets
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const int sx = (c & 1) + ex, sy = (c >> 1) + ey;  // source, relative to the first child
        if (sx >= -2 && sx <= 3 && sy >= -2 && sy <= 3) acc[c] = cfma(t, sv[sy + 2][sx + 2], acc[c]);
      }
    }
}

// The same sums source by source: each of the 36 sources is read once and
// added into every child it is a far-field source of, with that offset's
// operator (read from LDS per use): registers for 4 accumulators instead of
// the 36-source neighbourhood.
template <typename T>
__device__ __forceinline__ void par_translate_stream(const cplx<T> (*src)[kParKC], const cplx<T> (*tr)[kParKC],
                                                     int lx0, int ly0, int kk, cplx<T> (&acc)[4]) {
#pragma unroll
  for (int c = 0; c < 4; c++) acc[c] = mkc<T>(0, 0);
#pragma unroll
  for (int sy = -2; sy <= 3; sy++)
#pragma unroll
    for (int sx = -2; sx <= 3; sx++) {
      const cplx<T> v = src[(ly0 + sy + 2) * kParSB + lx0 + sx + 2][kk];
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const int ex = sx - (c & 1), ey = sy - (c >> 1);  // offset of the source from child c
        if (ex < -3 || ex > 3 || ey < -3 || ey > 3) continue;
        if (ex >= -1 && ex <= 1 && ey >= -1 && ey <= 1) continue;  // near offsets
        acc[c] = cfma(tr[7 * (ey + 3) + ex + 3][kk], v, acc[c]);
      }
    }
}

// (at least 2 waves per SIMD: the FP64 instantiation otherwise takes 256+ registers
// and runs one wave per SIMD; measured 2.00 -> 1.52 ms per FP64 launch set)
template <typename T>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2))) k_m2l_par(const cplx<T>* __restrict__ agm, cplx<T>* __restrict__ agl,
                                                    long st, const cplx<T>* __restrict__ trans, int lev, int S,
                                                    const int* __restrict__ act) {
  __shared__ cplx<T> src[kParSB * kParSB][kParKC];
  __shared__ cplx<T> tr[49][kParKC];
  const int rhs = blockIdx.y;
  if (rhs_off(act, rhs)) return;
  const ParBlock B(S);
  m2l_stage<T, kParTS, kParKC>(src, tr, agm + rhs * st, trans, B.X0, B.Y0, B.k0, 1 << lev, S);
  __syncthreads();
  const int kk = threadIdx.x % kParKC, pp = threadIdx.x / kParKC;
  const int lx0 = 2 * (pp & 3), ly0 = 2 * (pp >> 2);  // first child of the parent, tile coordinates
  cplx<T> acc[4];
  if constexpr (sizeof(T) == 8) {
    par_translate_stream<T>(src, tr, lx0, ly0, kk, acc);
  } else {
    cplx<T> sv[6][6];
    par_load<T>(src, lx0, ly0, kk, sv);
    par_translate<T>(tr, kk, sv, acc);
  }
  if (B.k0 + kk < S) {
    cplx<T>* out = agl + rhs * st + (long)(B.m0 + 4 * morton2((unsigned)(pp & 3), (unsigned)(pp >> 2))) * S;
#pragma unroll
    for (int c = 0; c < 4; c++) out[(long)c * S + B.k0 + kk] = acc[c];
  }
}

// Leaf M2L fused with the last L2L (parent level lev-1 -> leaf level lev):
// the k_m2l_par workgroup (4x4 parents, 16 samples = one 16-row block of the
// L2L band matrix) also interpolates its parents' local expansions to the
// children's samples on the matrix cores (as k_interp_band in L2L mode,
// columns = the 64 children, window = the row block's KB parent samples
// from s0v[chunk], mod Sp) and writes  local = translations + interpolation
// once.  The separate L2L pass would read the parents again and read and
// rewrite every leaf local after the M2L wrote it.  Dynamic LDS: the parent
// window, max(16 KB, 1024) elements (later the 64 x 16 interpolated
// locals), and the child shifts on it, 4 x KB.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_m2l_leaf(const cplx<T>* __restrict__ agm, cplx<T>* __restrict__ agl,
                                                     long st, const cplx<T>* __restrict__ trans, int lev, int S,
                                                     const cplx<T>* __restrict__ par, long ps, int Sp,
                                                     const T* __restrict__ AB, const int* __restrict__ s0v, int KB,
                                                     const cplx<T>* __restrict__ shift, const int* __restrict__ act) {
  // FP64: the translation streams its sources (par_translate_stream); the
  // neighbourhood in registers would take 256 registers, one wave per SIMD
  constexpr bool STREAM_LEAF = sizeof(T) == 8;
  typedef typename Mfma<T>::v4 v4;
  __shared__ cplx<T> src[kParSB * kParSB][kParKC];
  __shared__ cplx<T> tr[49][kParKC];
  extern __shared__ char smem_raw[];
  cplx<T>* pw = reinterpret_cast<cplx<T>*>(smem_raw);  // [16 parents][KB]
  cplx<T>* sw = pw + max(16 * KB, 64 * kParKC);         // [4 children][KB]
  const int rhs = blockIdx.y;
  if (rhs_off(act, rhs)) return;
  const ParBlock B(S);
  // the band rows of the window (the interpolation's A operand; no LDS
  // needed), issued before the staging so their latency overlaps it
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  constexpr int KQ = 24;  // windows of up to 96 samples (the L2L band limit)
  const int nkq = KB >> 2;
  T a[KQ];
  {
    const T* ab = AB + (long)B.chunk * KB * 16;
#pragma unroll
    for (int q = 0; q < KQ; q++) a[q] = (q < nkq) ? ab[(4 * q + lk) * 16 + li] : (T)0;
  }
  m2l_stage<T, kParTS, kParKC>(src, tr, agm + rhs * st, trans, B.X0, B.Y0, B.k0, 1 << lev, S);
  const int s0 = s0v[B.chunk];
  const cplx<T>* pr = par + rhs * ps + (long)(B.m0 >> 2) * Sp;  // the tile's 16 parents, Morton order
  // window element e = p * KB + jj (parent p, or child position p for the
  // shifts), parent sample (s0 + jj) mod Sp; (p, jj) stepped, not divided
  {
    const int n = 20 * KB, dp = kBlock / KB, dq = kBlock - dp * KB;
    int p = threadIdx.x / KB, jj = threadIdx.x - p * KB;
    for (int e0 = threadIdx.x; e0 < n; e0 += 4 * kBlock) {
      cplx<T> v[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int e = e0 + u * kBlock;
        int j = s0 + jj;
        if (j >= Sp) j -= Sp;
        // unconditional loads (past n: a clamped, unused shift element), so v stays in registers
        v[u] = e < 16 * KB ? pr[(long)p * Sp + j] : shift[min(p - 16, 3) * Sp + j];
        jj += dq;
        p += dp;
        if (jj >= KB) {
          jj -= KB;
          p++;
        }
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int e = e0 + u * kBlock;
        if (e < 16 * KB)
          pw[e] = v[u];
        else if (e < n)
          sw[e - 16 * KB] = v[u];
      }
    }
  }
  __syncthreads();
  const int tid = threadIdx.x, kk = tid % kParKC, pp = tid / kParKC;
  // interpolation first (the 36 sources of the translation phase are loaded
  // only after it, so the two phases do not hold registers at the same time):
  // wave w owns columns 16w .. 16w+15 (children 16w + i, i.e.
  // parent 4w + i/4, child position i%4); lane (li, lk) feeds column li at
  // window sample 4q + lk and receives rows Mfma::row(lk, q) at sample li
  const int w = tid >> 6;
  const cplx<T>* pcol = pw + (4 * w + (li >> 2)) * KB;
  const cplx<T>* scol = sw + (li & 3) * KB;
  v4 cr = v4{0, 0, 0, 0}, ci = v4{0, 0, 0, 0};
#pragma unroll
  for (int q = 0; q < KQ; q++) {
    if (q < nkq) {
      const int jj = 4 * q + lk;
      const cplx<T> bv = cmul(scol[jj], pcol[jj]);
      cr = Mfma<T>::mac(bv.x, a[q], cr);
      ci = Mfma<T>::mac(bv.y, a[q], ci);
    }
  }
  __syncthreads();  // the parent window is no longer read: it takes the result
  cplx<T>* ll = pw;  // [64 children][16 samples]
#pragma unroll
  for (int q = 0; q < 4; q++) ll[(16 * w + Mfma<T>::row(lk, q)) * kParKC + li] = mkc<T>(cr[q], ci[q]);
  __syncthreads();
  cplx<T> acc[4];
  if constexpr (STREAM_LEAF) {
    par_translate_stream<T>(src, tr, 2 * (pp & 3), 2 * (pp >> 2), kk, acc);
  } else {
    cplx<T> sv[6][6];
    par_load<T>(src, 2 * (pp & 3), 2 * (pp >> 2), kk, sv);
    par_translate<T>(tr, kk, sv, acc);
  }
  const unsigned pm = morton2((unsigned)(pp & 3), (unsigned)(pp >> 2));
  if (B.k0 + kk < S) {
    cplx<T>* out = agl + rhs * st + (long)(B.m0 + 4 * pm) * S;
#pragma unroll
    for (int c = 0; c < 4; c++) out[(long)c * S + B.k0 + kk] = cadd(acc[c], ll[(4 * pm + c) * kParKC + kk]);
  }
}

// P2M on the matrix cores: agm[box][k] = sum_n coeffT[n][k] y[box*64+n] for a
// tile of NB boxes.  The work is cut into units of 16 rows k x 16 boxes, dealt
// round-robin to the 4 waves: with S = 68 (5 row blocks) and 64-box tiles
// every wave gets 5 units (whole row blocks per wave would leave three waves
// idle half the time).  PREP (FP64 only): y holds x and the contrast product
// prep(x) (g x forward, conj(x) adjoint) is formed while loading it.
// With S <= kP2mRows (at most 5 row blocks) the units' results stay in
// registers until every wave is done, then go through LDS so that the tile's
// contiguous output range [b0 S, (b0 + NB) S) is written with whole,
// coalesced lines (stored from the MFMA layout, a box's row blocks came from
// different waves and partial lines reached HBM separately).
constexpr int kP2mRows = 80;
template <typename T, typename TI, int NB, bool PREP = false>
__global__ void __launch_bounds__(kBlock) k_p2m_mfma(const cplx<TI>* __restrict__ y, long ys,
                                                     cplx<T>* __restrict__ agm, long as,
                                                     const cplx<T>* __restrict__ coeffT, int S,
                                                     const int* __restrict__ act, const zd* __restrict__ g = nullptr,
                                                     long gs = 0, int herm = 0) {
  typedef typename Mfma<T>::v4 v4;
  constexpr int NN = NB / 16;
  constexpr int MAXU = (5 * NN + 3) / 4;  // units per wave with 5 row blocks
  __shared__ cplx<T> tile[NB * kP2mRows];  // input rows of 65, later the output rows of S
  const int rhs = blockIdx.y;
  if (rhs_off(act, rhs)) return;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const long b0 = (long)xcd_block() * NB;
  const cplx<TI>* yr = y + rhs * ys + b0 * 64;
  // loads issued in groups ahead of their LDS stores
  stage<sizeof(T) == 4 ? 8 : 4, cplx<TI>>(
      NB * 64,
      [&](int idx) {
        cplx<TI> v = yr[idx];
        if constexpr (PREP) v = prep_val(v, g, rhs * gs + b0 * 64 + idx, herm != 0);
        return v;
      },
      [&](int idx, cplx<TI> v) { tile[(idx >> 6) * 65 + (idx & 63)] = ccast<T>(v); });
  __syncthreads();
  const int nrb = (S + 15) / 16, nunit = nrb * NN;
  const bool staged = S <= kP2mRows;
  cplx<T>* out = agm + rhs * as;
  cplx<T> res[MAXU][4];
  // one unit: rows 16 rb .. +15 of the 16 boxes 16 j .. +15
  auto unit = [&](int u, cplx<T>(&o)[4]) {
    const int rb = u / NN, j = u - rb * NN;
    v4 p1 = v4{0, 0, 0, 0}, p2 = v4{0, 0, 0, 0}, p3 = v4{0, 0, 0, 0};
    const int krow = 16 * rb + li;
    const cplx<T>* trow = tile + (16 * j + li) * 65;
    // the coefficients of PRE k-steps are loaded together ahead of their MFMAs
    constexpr int PRE = sizeof(T) == 4 ? 8 : 4;
#pragma unroll 1
    for (int k0 = 0; k0 < 16; k0 += PRE) {
      cplx<T> a[PRE];
#pragma unroll
      for (int v = 0; v < PRE; v++) a[v] = (krow < S) ? coeffT[(4 * (k0 + v) + lk) * S + krow] : mkc<T>(0, 0);
#pragma unroll
      for (int v = 0; v < PRE; v++) {
        const T asum = a[v].x + a[v].y;
        Mfma<T>::cmac3(trow[4 * (k0 + v) + lk], a[v], asum, p1, p2, p3);
      }
    }
#pragma unroll
    for (int q = 0; q < 4; q++) o[q] = mkc<T>(p1[q] - p2[q], p3[q] - p1[q] - p2[q]);
  };
  if (staged) {
    // the wave's units as one software pipeline over (unit, k-batch) steps:
    // the coefficients of the next step are loaded while the current step's
    // MFMAs run (coeffT comes from L2; one step waiting for its own loads per
    // batch was the kernel's critical path)
    constexpr int PRE = sizeof(T) == 4 ? 8 : 4;
    constexpr int NBAT = 16 / PRE;  // k-batches per unit
    cplx<T> ab[2][PRE];
    auto load_step = [&](int c, int b, cplx<T>(&a)[PRE]) {
      const int u = w + 4 * c, krow = 16 * (u / NN) + li;
#pragma unroll
      for (int v = 0; v < PRE; v++)
        a[v] = (u < nunit && krow < S) ? coeffT[(4 * (b * PRE + v) + lk) * S + krow] : mkc<T>(0, 0);
    };
    load_step(0, 0, ab[0]);
#pragma unroll
    for (int c = 0; c < MAXU; c++) {
      const int u = w + 4 * c;
      const int j = u - (u / NN) * NN;
      const cplx<T>* trow = tile + (16 * j + li) * 65;
      v4 p1 = v4{0, 0, 0, 0}, p2 = v4{0, 0, 0, 0}, p3 = v4{0, 0, 0, 0};
#pragma unroll
      for (int b = 0; b < NBAT; b++) {
        const int i = c * NBAT + b;
        if (i + 1 < MAXU * NBAT) load_step((i + 1) / NBAT, (i + 1) % NBAT, ab[(i + 1) & 1]);
        if (u < nunit) {
#pragma unroll
          for (int v = 0; v < PRE; v++) {
            const cplx<T> a = ab[i & 1][v];
            Mfma<T>::cmac3(trow[4 * (b * PRE + v) + lk], a, a.x + a.y, p1, p2, p3);
          }
        }
      }
#pragma unroll
      for (int q = 0; q < 4; q++) res[c][q] = mkc<T>(p1[q] - p2[q], p3[q] - p1[q] - p2[q]);
    }
    __syncthreads();  // the input tile is no longer read: it takes the output
#pragma unroll
    for (int c = 0; c < MAXU; c++) {
      const int u = w + 4 * c;
      const int rb = u / NN, j = u - rb * NN, krow = 16 * rb + li;
      if (u < nunit && krow < S) {
#pragma unroll
        for (int q = 0; q < 4; q++) tile[(16 * j + Mfma<T>::row(lk, q)) * S + krow] = res[c][q];
      }
    }
    __syncthreads();
    cplx<T>* ob = out + b0 * S;
    for (int e = tid; e < NB * S; e += kBlock) ob[e] = tile[e];
  } else {
    for (int u = w; u < nunit; u += 4) {
      cplx<T> o[4];
      unit(u, o);
      const int rb = u / NN, j = u - rb * NN, krow = 16 * rb + li;
      if (krow < S) {
#pragma unroll
        for (int q = 0; q < 4; q++) out[(b0 + 16 * j + Mfma<T>::row(lk, q)) * S + krow] = o[q];
      }
    }
  }
}

// receiver values: out[m] = sum_t wT[t][m] agm0[(i0[m]+t) mod S0]
template <typename T>
__global__ void __launch_bounds__(kBlock) k_rx_interp(const cplx<T>* __restrict__ agm0, long as, zd* __restrict__ out,
                                                      long os, const T* __restrict__ wT, const int* __restrict__ i0,
                                                      int S0, int R, int ntap) {
  int rhs = blockIdx.y;
  int m = blockIdx.x * kBlock + threadIdx.x;
  if (m >= R) return;
  const cplx<T>* a = agm0 + rhs * as;
  cplx<T> acc = mkc<T>(0, 0);
  int j = i0[m];
  for (int t = 0; t < ntap; t++) {
    T w = wT[t * R + m];
    acc.x += w * a[j].x;
    acc.y += w * a[j].y;
    j = (j + 1 == S0) ? 0 : j + 1;
  }
  out[rhs * os + m] = ccast<double>(acc);
}

// root locals from receiver values: agl0[k] = (sum_t aT[t][k] in[aI[t][k]]) * (-4i) * sc
template <typename T>
__global__ void __launch_bounds__(kBlock) k_rx_anterp(const zd* __restrict__ in, long is, cplx<T>* __restrict__ agl0,
                                                      long as, const T* __restrict__ aT, const int* __restrict__ aI,
                                                      int S0, int ntap, double sc) {
  int rhs = blockIdx.y;
  int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= S0) return;
  const zd* v = in + rhs * is;
  zd acc = mkc<double>(0, 0);
  for (int t = 0; t < ntap; t++) {
    double w = (double)aT[t * S0 + k];
    zd u = v[aI[t * S0 + k]];
    acc.x += w * u.x;
    acc.y += w * u.y;
  }
  zd r = cmul(acc, mkc<double>(0.0, -4.0));
  r = mkc<double>(r.x * sc, r.y * sc);
  agl0[rhs * as + k] = ccast<T>(r);
}

}  // namespace isg
