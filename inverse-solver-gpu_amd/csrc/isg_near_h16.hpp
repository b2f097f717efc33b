// isg_near_h16.hpp -- near field + L2P (k_near_h16) and P2M (k_p2m_h16) of the
// FP32 operator on gfx950's FP16 matrix cores with split operands
// (v_mfma_f32_16x16x32_f16).
//
// Every FP32 operand x is carried as two FP16 numbers, hi = rne16(2^e x) and
// lo = rne16(2^e x - hi), with a power-of-two scale 2^e chosen so that
// |2^e x| < 2^14 -- per table, and per workgroup and right-hand side for the
// data.  hi + lo holds x to 2^-22 relative (elements below 2^-16 of the
// largest one, whose lo is subnormal, to 2^-38 of the largest, absolutely).  A product x a is formed
// as hi.hi + hi.lo + lo.hi in FP32 accumulators (the dropped lo.lo term is
// below 2^-22 relative): three FP16 MFMAs, each 16 times the FP32 rate, for
// one FP32 MFMA's worth of products.  Complex products take the four-product
// form (cr += dr ar - di ai, ci += dr ai + di ar): 12 MFMAs per 16 x 16 x 32
// block.  The scales are exact powers of two and are divided out of the FP32
// accumulators at the end.
//
// Workgroup: 8 x 8 Morton-consecutive leaf boxes (64) of two right-hand sides
// = 128 MFMA rows, so each table fragment read from L2 feeds 8 row blocks
// (the table, 8 B per complex entry in its four FP16 parts, is what limits a
// split kernel: the matrix cores consume it ~5 times faster than the FP32
// kernel does, which reads the same 8 B per entry for 16 or 32 rows).  Wave w
// owns right-hand side w / 2 of the pair (its 4 row blocks of 16 boxes) and
// output pixels 32 (w % 2) .. +31 (2 column blocks): every A fragment read
// from LDS feeds 2 column blocks, every B fragment 4 row blocks.
//
// Reference counterpart: the near-field and L2P stages of matvec_gpu.cu
// (kernels.cu:32-81, :292-319); the result equals the FP32 kernel's to the
// split's 2^-22.
#pragma once

#include "isg_kernels.hpp"

namespace isg {

typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef unsigned v4u __attribute__((ext_vector_type(4)));

// Tuning value of the near stage that selects this kernel (FP32 operator).
constexpr int kNearSplit = 128;
// Halo of the 8 x 8 box tile: 10 x 10 boxes at a row pitch of 12 positions,
// so the 16 boxes of a row block (a 4 x 4 block) read by 16 lanes fall in 16
// different 16-byte bank groups.  Positions (and the boxes of the L2P chunk)
// are stored XOR 2 g for k group g, so that the four groups of one box, which
// consecutive lanes write, also fall in different bank groups (a bijection on
// the low bits: readers of one group keep distinct banks; 119 ^ 6 < 120).
constexpr int kH16Pitch = 12;
constexpr int kH16Pos = kH16Pitch * 10;
// LDS (bytes): half of the halo (32 of the 64 pixels of each box) as
// [rhs 2][part 4][k group 4][position 120] x 8 FP16; reused for a 32-local
// chunk of the L2P ([rhs][part][k group][box 64] x 8 FP16) and for one
// right-hand side's 64 x 64 complex output tile.
constexpr size_t kH16Lds = (size_t)2 * 4 * 4 * kH16Pos * 16;
static_assert(kH16Lds >= (size_t)64 * 64 * 8 && kH16Lds >= (size_t)2 * 4 * 4 * 64 * 16, "LDS reuse");

// Power-of-two exponent e with |2^e m| < 2^14 for the largest magnitude m
// (0 for m = 0 or m not finite: Inf / NaN then propagate as in FP32).
__host__ __device__ inline int h16_scale_exp(float m) {
  if (!(m > 0.f) || !(m <= 3.0e38f)) return 0;
  int e;
  (void)frexpf(m, &e);  // m = f 2^e, f in [0.5, 1)
  e = 14 - e;
  return e > 126 ? 126 : (e < -126 ? -126 : e);
}

// element e of hi and lo: the split of v
__device__ __forceinline__ void h16_split(float v, v8h& hi, v8h& lo, int e) {
  const _Float16 h = (_Float16)v;
  hi[e] = h;
  lo[e] = (_Float16)(v - (float)h);  // exact difference: at most 13 significant bits
}

// Four complex values (two float4) scaled by s and split into the four parts
// (re hi, re lo, im hi, im lo), elements e0 .. e0 + 3 of each part.
__device__ __forceinline__ void h16_split4(float4 a, float4 b, float s, v8h (&ph)[4], int e0) {
  h16_split(a.x * s, ph[0], ph[1], e0);
  h16_split(a.y * s, ph[2], ph[3], e0);
  h16_split(a.z * s, ph[0], ph[1], e0 + 1);
  h16_split(a.w * s, ph[2], ph[3], e0 + 1);
  h16_split(b.x * s, ph[0], ph[1], e0 + 2);
  h16_split(b.y * s, ph[2], ph[3], e0 + 2);
  h16_split(b.z * s, ph[0], ph[1], e0 + 3);
  h16_split(b.w * s, ph[2], ph[3], e0 + 3);
}

__device__ __forceinline__ float h16_absmax(float4 v) {
  return fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w)));
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

// Workgroup maxima of N values; every thread gets all of them.
template <int N>
__device__ __forceinline__ void block_max(float (&m)[N], float (*red)[kBlock / 64]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < N; q++) m[q] = wave_max(m[q]);
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < N; q++) red[q][w] = m[q];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < N; q++) {
    float v = red[q][0];
#pragma unroll
    for (int i = 1; i < kBlock / 64; i++) v = fmaxf(v, red[q][i]);
    m[q] = v;
  }
}

// a[q] for a runtime q without indexing a private array (which would live in scratch)
template <typename T>
__device__ __forceinline__ T sel2(const T (&a)[2], int q) {
  return q ? a[1] : a[0];
}

__device__ __forceinline__ v8h h16_neg(v8h v) {
  return __builtin_bit_cast(v8h, __builtin_bit_cast(v4u, v) ^ 0x80008000u);
}

// cr + i ci += (a0 + a1 + i (a2 + a3)) (b0 + b1 + i (b2 + b3)) without the
// lo.lo terms; a = data parts (re hi, re lo, im hi, im lo), b = table parts,
// nb = -b2, -b3.  The two accumulators alternate so that consecutive MFMAs
// are independent.
__device__ __forceinline__ void h16_cmac(const v8h (&a)[4], const v8h (&b)[4], const v8h (&nb)[2], v4f& cr,
                                         v4f& ci) {
  cr = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[0], b[0], cr, 0, 0, 0);
  ci = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[0], b[2], ci, 0, 0, 0);
  cr = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[0], b[1], cr, 0, 0, 0);
  ci = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[0], b[3], ci, 0, 0, 0);
  cr = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[1], b[0], cr, 0, 0, 0);
  ci = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[1], b[2], ci, 0, 0, 0);
  cr = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[2], nb[0], cr, 0, 0, 0);
  ci = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[2], b[0], ci, 0, 0, 0);
  cr = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[2], nb[1], cr, 0, 0, 0);
  ci = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[2], b[1], ci, 0, 0, 0);
  cr = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[3], nb[0], cr, 0, 0, 0);
  ci = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[3], b[0], ci, 0, 0, 0);
}

// Split-FP16 table: out[b][part][g][m] (m < ncolp) = 8 FP16 parts of
// 2^ex src[(b / per) bstride + k srow + m], k = 32 (b % per) + 8 g + e,
// e = 0..7 (zero for k >= kmax or m >= ncol); one thread per (b, g, m).
static __global__ void __launch_bounds__(256) k_split16(const zf* __restrict__ src, int per, long bstride, int kmax,
                                                        int srow, int ncol, int ncolp, int ex, int nblk,
                                                        v8h* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nblk * 4 * ncolp) return;
  const int m = t % ncolp, g = (t / ncolp) & 3, b = t / (4 * ncolp);
  const float s = ldexpf(1.f, ex);
  v8h ph[4];
#pragma unroll
  for (int e = 0; e < 8; e++) {
    const int k = 32 * (b % per) + 8 * g + e;
    const zf v = (k < kmax && m < ncol) ? src[(long)(b / per) * bstride + (long)k * srow + m] : mkc<float>(0.f, 0.f);
    h16_split(v.x * s, ph[0], ph[1], e);
    h16_split(v.y * s, ph[2], ph[3], e);
  }
#pragma unroll
  for (int P = 0; P < 4; P++) out[((long)(b * 4 + P) * 4 + g) * ncolp + m] = ph[P];
}

// As h16_cmac with the negation on the data side: na = -a2, -a3 (the table
// fragments are then used as loaded; fewer live registers when one table
// fragment serves several data fragments one after another).
__device__ __forceinline__ void h16_cmac_na(const v8h (&a)[4], const v8h (&na)[2], const v8h (&b)[4], v4f& cr,
                                            v4f& ci) {
  cr = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[0], b[0], cr, 0, 0, 0);
  ci = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[0], b[2], ci, 0, 0, 0);
  cr = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[0], b[1], cr, 0, 0, 0);
  ci = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[0], b[3], ci, 0, 0, 0);
  cr = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[1], b[0], cr, 0, 0, 0);
  ci = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[1], b[2], ci, 0, 0, 0);
  cr = __builtin_amdgcn_mfma_f32_16x16x32_f16(na[0], b[2], cr, 0, 0, 0);
  ci = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[2], b[0], ci, 0, 0, 0);
  cr = __builtin_amdgcn_mfma_f32_16x16x32_f16(na[0], b[3], cr, 0, 0, 0);
  ci = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[2], b[1], ci, 0, 0, 0);
  cr = __builtin_amdgcn_mfma_f32_16x16x32_f16(na[1], b[2], cr, 0, 0, 0);
  ci = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[3], b[0], ci, 0, 0, 0);
}

// Split tables, built once per operator (k_split16):
//   tab[s][kc][part][g][m] (8 FP16 each): 2^et near[s][32 kc + 8 g + e][m], e = 0..7
//   bas[kc][part][g][m]    (8 FP16 each): 2^eb basisT[32 kc + 8 g + e][m] (0 for k >= S)
// Grid (B / 64, ceil(nrhs / 2)), kBlock threads, kH16Lds bytes of dynamic LDS.
template <int POST>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2)))
k_near_h16(const zf* __restrict__ y, long ys, const zf* __restrict__ loc, long ls, zf* __restrict__ r, long rs,
           const v8h* __restrict__ tab, int et, const v8h* __restrict__ bas, int eb, int S, int nbs, int nrhs,
           const int* __restrict__ act, NearPost<float> post) {
  extern __shared__ __attribute__((aligned(16))) char smem_raw[];
  v8h* hal = reinterpret_cast<v8h*>(smem_raw);
  __shared__ int hoff[100];
  __shared__ float red[4][kBlock / 64];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int rw = w >> 1, mb0 = 2 * (w & 1);
  const int rhs[2] = {2 * (int)blockIdx.y, 2 * (int)blockIdx.y + 1};
  bool on[2];
#pragma unroll
  for (int q = 0; q < 2; q++) on[q] = rhs[q] < nrhs && !rhs_off(act, rhs[q]);
  if (!on[0] && !on[1]) return;
  const bool mine = sel2(on, rw);  // wave-uniform
  // a right-hand side that loads may use when its own is off (never read then)
  const int rv[2] = {on[0] ? rhs[0] : rhs[1], on[1] ? rhs[1] : rhs[0]};
  const long b0 = (long)xcd_block() * 64;
  const int X0 = (int)compact2((unsigned)b0), Y0 = (int)compact2((unsigned)b0 >> 1);
  for (int h = tid; h < 100; h += kBlock) {
    const int sx = X0 - 1 + h % 10, sy = Y0 - 1 + h / 10;
    hoff[h] = (sx >= 0 && sy >= 0 && sx < nbs && sy < nbs) ? (int)morton2(sx, sy) * 64 : -1;
  }
  __syncthreads();

  // data scale of each right-hand side: largest |re|, |im| over the halo,
  // from one pass whose loads are all issued before the first comparison
  // (the accumulators are not live yet)
  int es[2], el[2];
  {
    constexpr int NU = (3200 + kBlock - 1) / kBlock;  // 100 boxes x 32 float4 per right-hand side
    float m[2] = {0.f, 0.f};
    float4 v[2][NU];
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const float4* yq = reinterpret_cast<const float4*>(y + (long)sel2(rv, q) * ys);
#pragma unroll
      for (int u = 0; u < NU; u++) {
        const int e = tid + u * kBlock, ec = e < 3200 ? e : 3199;
        const int off = hoff[ec >> 5];
        v[q][u] = yq[((off < 0 ? 0 : off) >> 1) + (ec & 31)];
        if (!(sel2(on, q) && e < 3200 && off >= 0)) v[q][u] = make_float4(0.f, 0.f, 0.f, 0.f);
      }
    }
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
      for (int u = 0; u < NU; u++) m[q] = fmaxf(m[q], h16_absmax(v[q][u]));
    block_max<2>(m, red);
#pragma unroll
    for (int q = 0; q < 2; q++) es[q] = h16_scale_exp(m[q]);
  }

  v4f cr[4][2], ci[4][2];
#pragma unroll
  for (int j = 0; j < 4; j++)
#pragma unroll
    for (int c = 0; c < 2; c++) {
      cr[j][c] = v4f{0, 0, 0, 0};
      ci[j][c] = v4f{0, 0, 0, 0};
    }
  // halo position of box 16 j + li (this lane's A row) for the centre slot
  int hc[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const unsigned c = (unsigned)(16 * j + li);
    hc[j] = ((int)compact2(c >> 1) + 1) * kH16Pitch + (int)compact2(c) + 1;
  }

  // ---- near field: two halves of the source pixels (k = 32 kc .. +31)
  // B fragments of slot s (this lane's 8 k of column li, for both column
  // blocks); the next slot's are loaded while the current slot's MFMAs run
  const v8h* tl = tab + lk * 64 + 16 * mb0 + li;  // + (s * 2 + kc) * 1024 + P * 256 + 16 c
  auto load_b = [&](v8h (&b)[2][4], const v8h* t) {
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
      for (int P = 0; P < 4; P++) b[c][P] = t[P * 256 + 16 * c];
  };
  v8h bn[2][4];
#pragma unroll 1
  for (int kc = 0; kc < 2; kc++) {
    __syncthreads();  // the previous half is no longer read
    if (mine) load_b(bn, tl + (long)kc * 1024);  // in flight during the staging
    {
      // units (q, h, g) of 8 source pixels 32 kc + 8 g .. +7 of halo box h
      float4 v[4][4];
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const int u = tid + t * kBlock, uc = u < 800 ? u : 799;
        const int q = uc / 400, h = (uc >> 2) % 100, g = uc & 3;
        const int off = hoff[h];
        const float4* src = reinterpret_cast<const float4*>(y + (long)sel2(rv, q) * ys + (off < 0 ? 0 : off) + 32 * kc + 8 * g);
#pragma unroll
        for (int c = 0; c < 4; c++) v[t][c] = src[c];
      }
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const int u = tid + t * kBlock;
        if (u < 800) {
          const int q = u / 400, h = (u >> 2) % 100, g = u & 3;
          const float s = ldexpf(1.f, sel2(es, q));
          // zero outside the grid and for a right-hand side that is off (its loads used stand-in addresses)
          const bool ok = sel2(on, q) && hoff[h] >= 0;
          const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
          v8h ph[4];
          h16_split4(ok ? v[t][0] : z, ok ? v[t][1] : z, s, ph, 0);
          h16_split4(ok ? v[t][2] : z, ok ? v[t][3] : z, s, ph, 4);
          const int p = (h / 10) * kH16Pitch + h % 10;
#pragma unroll
          for (int P = 0; P < 4; P++) hal[((q * 4 + P) * 4 + g) * kH16Pos + (p ^ (2 * g))] = ph[P];
        }
      }
    }
    __syncthreads();
    if (mine) {
      const v8h* hb = hal + (rw * 16 + lk) * kH16Pos;  // + P * 4 * kH16Pos + position
#pragma unroll 1
      for (int s = 0; s < 9; s++) {
        const int off = (s / 3 - 1) * kH16Pitch + (s % 3 - 1);
        v8h b[2][4], nb[2][2];
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
          for (int P = 0; P < 4; P++) b[c][P] = bn[c][P];
        if (s + 1 < 9) load_b(bn, tl + (long)((s + 1) * 2 + kc) * 1024);
#pragma unroll
        for (int c = 0; c < 2; c++) {
          nb[c][0] = h16_neg(b[c][2]);
          nb[c][1] = h16_neg(b[c][3]);
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const int p = (hc[j] + off) ^ (2 * lk);
          v8h a[4];
#pragma unroll
          for (int P = 0; P < 4; P++) a[P] = hb[P * 4 * kH16Pos + p];
#pragma unroll
          for (int c = 0; c < 2; c++) h16_cmac(a, b[c], nb[c], cr[j][c], ci[j][c]);
        }
      }
    }
  }

  // ---- L2P: locals of the tile's boxes, in chunks of 32; units (q, box, g):
  // locals 32 kc + 8 g .. +7 of box b0 + box, loaded a chunk ahead
  float4 lv[2][4];
  auto load_loc = [&](int kc) {
#pragma unroll
    for (int t = 0; t < 2; t++) {
      const int u = tid + t * kBlock;
      const int q = u >> 8, bx = (u >> 2) & 63, g = u & 3;
      const zf* lb = loc + (long)sel2(rv, q) * ls + (b0 + bx) * S;
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const int k = 32 * kc + 8 * g + 2 * c;
        lv[t][c] = *reinterpret_cast<const float4*>(lb + (k < S ? k : S - 2));
      }
    }
  };
  load_loc(0);
  // locals scale of each right-hand side, read after the near field so that the
  // staging below finds them in L2
  {
    float m[2] = {0.f, 0.f};
    const int n4 = 32 * S;  // 64 boxes x S locals = 32 S float4 (S even), per right-hand side
    constexpr int U = 9;
    for (int e0 = tid; e0 < 2 * n4; e0 += U * kBlock) {
      float4 t[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int e = e0 + u * kBlock, q = e >= n4 ? 1 : 0, ei = e - q * n4;
        const float4* lq = reinterpret_cast<const float4*>(loc + (long)sel2(rv, q) * ls + b0 * S);
        t[u] = lq[ei < n4 ? ei : n4 - 1];
        if (!(sel2(on, q) && e < 2 * n4)) t[u] = make_float4(0.f, 0.f, 0.f, 0.f);
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const float a = h16_absmax(t[u]);
        if (e0 + u * kBlock >= n4)
          m[1] = fmaxf(m[1], a);
        else
          m[0] = fmaxf(m[0], a);
      }
    }
    __syncthreads();  // red is read by the data scale above
    block_max<2>(m, red);
#pragma unroll
    for (int q = 0; q < 2; q++) el[q] = h16_scale_exp(m[q]);
  }
  if (mine) {
    // near accumulators carry 2^(es + et); bring them to the L2P's 2^(el + eb)
    const int n = (sel2(el, rw) + eb) - (sel2(es, rw) + et);
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int c = 0; c < 2; c++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
          cr[j][c][i] = ldexpf(cr[j][c][i], n);
          ci[j][c][i] = ldexpf(ci[j][c][i], n);
        }
  }
  const int nlc = (S + 31) / 32;
#pragma unroll 1
  for (int kc = 0; kc < nlc; kc++) {
    __syncthreads();
    if (mine) load_b(bn, bas + (long)kc * 1024 + lk * 64 + 16 * mb0 + li);  // in flight during the staging
    {
      // converts the chunk's locals, loaded during the previous chunk's MFMAs
#pragma unroll
      for (int t = 0; t < 2; t++) {
        const int u = tid + t * kBlock;
        const int q = u >> 8, bx = (u >> 2) & 63, g = u & 3;
        const float s = ldexpf(1.f, sel2(el, q));
        float4 v[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
          const int k = 32 * kc + 8 * g + 2 * c;  // pairs (k, k + 1) lie within S or beyond it together
          v[c] = (sel2(on, q) && k < S) ? lv[t][c] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        v8h ph[4];
        h16_split4(v[0], v[1], s, ph, 0);
        h16_split4(v[2], v[3], s, ph, 4);
#pragma unroll
        for (int P = 0; P < 4; P++) hal[((q * 4 + P) * 4 + g) * 64 + (bx ^ (2 * g))] = ph[P];
      }
    }
    __syncthreads();
    if (kc + 1 < nlc) load_loc(kc + 1);
    if (mine) {
      const v8h* hb = hal + (rw * 16 + lk) * 64;  // + P * 256 + box
      v8h nb[2][2];
#pragma unroll
      for (int c = 0; c < 2; c++) {
        nb[c][0] = h16_neg(bn[c][2]);
        nb[c][1] = h16_neg(bn[c][3]);
      }
#pragma unroll
      for (int j = 0; j < 4; j++) {
        v8h a[4];
#pragma unroll
        for (int P = 0; P < 4; P++) a[P] = hb[P * 256 + ((16 * j + li) ^ (2 * lk))];
#pragma unroll
        for (int c = 0; c < 2; c++) h16_cmac(a, bn[c], nb[c], cr[j][c], ci[j][c]);
      }
    }
  }

  // ---- output, one right-hand side at a time through LDS (coalesced stores
  // and, for a post step, coalesced loads); D[box][m]: box 16 j + 4 lk + i
  // (register i), pixel m = 16 (mb0 + c) + li
  zf* tile = reinterpret_cast<zf*>(smem_raw);
#pragma unroll 1
  for (int q = 0; q < 2; q++) {
    if (!sel2(on, q)) continue;  // uniform
    __syncthreads();
    if (rw == q) {
      const int n = -(sel2(el, q) + eb);
#pragma unroll
      for (int j = 0; j < 4; j++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
          for (int i = 0; i < 4; i++)
            tile[(16 * j + 4 * lk + i) * 64 + 16 * (mb0 + c) + li] =
                mkc<float>(ldexpf(cr[j][c][i], n), ldexpf(ci[j][c][i], n));
    }
    __syncthreads();
    const long o = (long)sel2(rhs, q) * rs + b0 * 64;
    if constexpr (POST == 0) {
#pragma unroll 4
      for (int e = tid; e < 64 * 64; e += kBlock) r[o + e] = tile[e];
    } else {
      const zd* gr = post.g + (long)sel2(rhs, q) * post.gs + b0 * 64;
      zd acc[2] = {mkc<double>(0, 0), mkc<double>(0, 0)};
      struct In {
        zf x, aux;  // aux: r~ (mode 1)
        zd g;
      };
      const bool herm = post.herm != 0;
      stage<2, In>(
          64 * 64,
          [&](int e) {
            In v;
            v.x = post.x[o + e];
            if (POST == 1) v.aux = post.rt[o + e];
            if (herm) v.g = gr[e];  // the forward post step does not use g
            return v;
          },
          [&](int e, In v) {
            const zf gt = tile[e];
            const zd gy = mkc<double>((double)gt.x, (double)gt.y);
            const zd xv = ccast<double>(v.x);
            const zd px = herm ? csub(xv, cconj(cmul(v.g, gy))) : csub(xv, gy);
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
          });
      block_store<2>(acc, post.part + (long)sel2(rhs, q) * kSlots * gridDim.x, gridDim.x, blockIdx.x);
    }
  }
}

// ---------------------------------------------------------------------------
// Split-FP16 near field + L2P with 128-box tiles of ONE right-hand side
// (tuning value kNearSplitWide): the same arithmetic and table as k_near_h16,
// the same 128 MFMA rows per table fragment, but a 16 x 8 box tile (halo
// 18 x 10 boxes: 1.41 x the tile's pixels staged instead of 1.56 x), no
// half-idle pairs when systems of a batch have converged, and 51 KB of LDS
// and at most 168 registers, so three workgroups share a CU (k_near_h16: two).
// Wave w owns row blocks 4 (w / 2) .. +3 (16 boxes each) and column blocks
// 2 (w % 2), +1.  Halo positions at a row pitch of 20 (= 4 mod 8, as 12 is),
// stored XOR 2 g as in k_near_h16.  Grid (B / 128, nrhs).
constexpr int kNearSplitWide = 256;
constexpr int kH16WPitch = 20;
constexpr int kH16WPos = kH16WPitch * 10;
constexpr size_t kH16WLds = (size_t)4 * 4 * kH16WPos * 16;
static_assert(kH16WLds >= (size_t)64 * 64 * 8 && kH16WLds >= (size_t)4 * 4 * 128 * 16, "LDS reuse");

template <int POST>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2)))
k_near_h16w(const zf* __restrict__ y, long ys, const zf* __restrict__ loc, long ls, zf* __restrict__ r, long rs,
            const v8h* __restrict__ tab, int et, const v8h* __restrict__ bas, int eb, int S, int nbs,
            const int* __restrict__ act, NearPost<float> post) {
  extern __shared__ __attribute__((aligned(16))) char smem_raw[];
  v8h* hal = reinterpret_cast<v8h*>(smem_raw);
  __shared__ int hoff[180];
  __shared__ float red[4][kBlock / 64];
  const int rhs = blockIdx.y;
  if (rhs_off(act, rhs)) return;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int rbh = w >> 1, mb0 = 2 * (w & 1);
  const long b0 = (long)xcd_block() * 128;
  const int X0 = (int)compact2((unsigned)b0), Y0 = (int)compact2((unsigned)b0 >> 1);
  for (int h = tid; h < 180; h += kBlock) {
    const int sx = X0 - 1 + h % 18, sy = Y0 - 1 + h / 18;
    hoff[h] = (sx >= 0 && sy >= 0 && sx < nbs && sy < nbs) ? (int)morton2(sx, sy) * 64 : -1;
  }
  __syncthreads();
  const zf* yr = y + (long)rhs * ys;

  // data scale: largest |re|, |im| over the halo, all loads issued first
  int es;
  {
    constexpr int NU = 12;  // 180 boxes x 32 float4 = 5760 in two batches of 12 per thread
    float m[1] = {0.f};
    const float4* yq = reinterpret_cast<const float4*>(yr);
#pragma unroll 1
    for (int e0 = tid; e0 < 5760; e0 += NU * kBlock) {
      float4 v[NU];
#pragma unroll
      for (int u = 0; u < NU; u++) {
        const int e = e0 + u * kBlock, ec = e < 5760 ? e : 5759;
        const int off = hoff[ec >> 5];
        v[u] = yq[((off < 0 ? 0 : off) >> 1) + (ec & 31)];
        if (!(e < 5760 && off >= 0)) v[u] = make_float4(0.f, 0.f, 0.f, 0.f);
      }
#pragma unroll
      for (int u = 0; u < NU; u++) m[0] = fmaxf(m[0], h16_absmax(v[u]));
    }
    block_max<1>(m, red);
    es = h16_scale_exp(m[0]);
  }

  v4f cr[4][2], ci[4][2];
#pragma unroll
  for (int j = 0; j < 4; j++)
#pragma unroll
    for (int c = 0; c < 2; c++) {
      cr[j][c] = v4f{0, 0, 0, 0};
      ci[j][c] = v4f{0, 0, 0, 0};
    }
  // halo position of box 16 (4 rbh + j) + li for the centre slot
  int hc[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const unsigned c = (unsigned)(16 * (4 * rbh + j) + li);
    hc[j] = ((int)compact2(c >> 1) + 1) * kH16WPitch + (int)compact2(c) + 1;
  }

  const v8h* tl = tab + lk * 64 + 16 * mb0 + li;  // + (s * 2 + kc) * 1024 + P * 256 + 16 c
  auto load_b = [&](v8h (&b)[2][4], const v8h* t) {
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
      for (int P = 0; P < 4; P++) b[c][P] = t[P * 256 + 16 * c];
  };
  const float sd = ldexpf(1.f, es);
  v8h bn[2][4];
#pragma unroll 1
  for (int kc = 0; kc < 2; kc++) {
    __syncthreads();  // the previous half is no longer read
    load_b(bn, tl + (long)kc * 1024);  // in flight during the staging
    {
      // units (h, g) of 8 source pixels 32 kc + 8 g .. +7 of halo box h
      float4 v[3][4];
#pragma unroll
      for (int t = 0; t < 3; t++) {
        const int u = tid + t * kBlock, uc = u < 720 ? u : 719;
        const int h = uc >> 2, g = uc & 3;
        const int off = hoff[h];
        const float4* src = reinterpret_cast<const float4*>(yr + (off < 0 ? 0 : off) + 32 * kc + 8 * g);
#pragma unroll
        for (int c = 0; c < 4; c++) v[t][c] = src[c];
      }
#pragma unroll
      for (int t = 0; t < 3; t++) {
        const int u = tid + t * kBlock;
        if (u < 720) {
          const int h = u >> 2, g = u & 3;
          const bool ok = hoff[h] >= 0;  // zero outside the grid
          const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
          v8h ph[4];
          h16_split4(ok ? v[t][0] : z, ok ? v[t][1] : z, sd, ph, 0);
          h16_split4(ok ? v[t][2] : z, ok ? v[t][3] : z, sd, ph, 4);
          const int p = (h / 18) * kH16WPitch + h % 18;
#pragma unroll
          for (int P = 0; P < 4; P++) hal[(P * 4 + g) * kH16WPos + (p ^ (2 * g))] = ph[P];
        }
      }
    }
    __syncthreads();
    const v8h* hb = hal + lk * kH16WPos;  // + P * 4 * kH16WPos + position
#pragma unroll 1
    for (int s = 0; s < 9; s++) {
      const int off = (s / 3 - 1) * kH16WPitch + (s % 3 - 1);
      v8h b[2][4];
#pragma unroll
      for (int c = 0; c < 2; c++)
#pragma unroll
        for (int P = 0; P < 4; P++) b[c][P] = bn[c][P];
      if (s + 1 < 9) load_b(bn, tl + (long)((s + 1) * 2 + kc) * 1024);  // the next slot's, during these MFMAs
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int p = (hc[j] + off) ^ (2 * lk);
        v8h a[4], na[2];
#pragma unroll
        for (int P = 0; P < 4; P++) a[P] = hb[P * 4 * kH16WPos + p];
        na[0] = h16_neg(a[2]);
        na[1] = h16_neg(a[3]);
#pragma unroll
        for (int c = 0; c < 2; c++) h16_cmac_na(a, na, b[c], cr[j][c], ci[j][c]);
      }
    }
  }

  // ---- L2P: locals of the tile's 128 boxes in chunks of 32; units (box, g):
  // locals 32 kc + 8 g .. +7 of box b0 + box, loaded a chunk ahead
  const zf* lr = loc + (long)rhs * ls + b0 * S;
  float4 lv[2][4];
  auto load_loc = [&](int kc) {
#pragma unroll
    for (int t = 0; t < 2; t++) {
      const int u = tid + t * kBlock;
      const int bx = u >> 2, g = u & 3;
      const zf* lb = lr + (long)bx * S;
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const int k = 32 * kc + 8 * g + 2 * c;
        lv[t][c] = *reinterpret_cast<const float4*>(lb + (k < S ? k : S - 2));
      }
    }
  };
  load_loc(0);
  int el;
  {
    float m[1] = {0.f};
    const int n4 = 64 * S;  // 128 boxes x S locals = 64 S float4 (S even)
    constexpr int U = 9;
    const float4* lq = reinterpret_cast<const float4*>(lr);
    for (int e0 = tid; e0 < n4; e0 += U * kBlock) {
      float4 t[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int e = e0 + u * kBlock;
        t[u] = lq[e < n4 ? e : n4 - 1];
      }
#pragma unroll
      for (int u = 0; u < U; u++) m[0] = fmaxf(m[0], h16_absmax(t[u]));  // clamped repeats change no maximum
    }
    __syncthreads();  // red is read by the data scale above
    block_max<1>(m, red);
    el = h16_scale_exp(m[0]);
  }
  {
    // near accumulators carry 2^(es + et); bring them to the L2P's 2^(el + eb)
    const int n = (el + eb) - (es + et);
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int c = 0; c < 2; c++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
          cr[j][c][i] = ldexpf(cr[j][c][i], n);
          ci[j][c][i] = ldexpf(ci[j][c][i], n);
        }
  }
  const float sl = ldexpf(1.f, el);
  const int nlc = (S + 31) / 32;
#pragma unroll 1
  for (int kc = 0; kc < nlc; kc++) {
    __syncthreads();
    v8h b[2][4];
    load_b(b, bas + (long)kc * 1024 + lk * 64 + 16 * mb0 + li);  // in flight during the staging
    {
#pragma unroll
      for (int t = 0; t < 2; t++) {
        const int u = tid + t * kBlock;
        const int bx = u >> 2, g = u & 3;
        float4 v[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
          const int k = 32 * kc + 8 * g + 2 * c;  // pairs (k, k + 1) lie within S or beyond it together
          v[c] = k < S ? lv[t][c] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        v8h ph[4];
        h16_split4(v[0], v[1], sl, ph, 0);
        h16_split4(v[2], v[3], sl, ph, 4);
#pragma unroll
        for (int P = 0; P < 4; P++) hal[(P * 4 + g) * 128 + (bx ^ (2 * g))] = ph[P];
      }
    }
    __syncthreads();
    if (kc + 1 < nlc) load_loc(kc + 1);
    const v8h* hb = hal + lk * 128;  // + P * 512 + box
#pragma unroll
    for (int j = 0; j < 4; j++) {
      v8h a[4], na[2];
#pragma unroll
      for (int P = 0; P < 4; P++) a[P] = hb[P * 512 + ((16 * (4 * rbh + j) + li) ^ (2 * lk))];
      na[0] = h16_neg(a[2]);
      na[1] = h16_neg(a[3]);
#pragma unroll
      for (int c = 0; c < 2; c++) h16_cmac_na(a, na, b[c], cr[j][c], ci[j][c]);
    }
  }

  // ---- output in two halves of 64 boxes through LDS; D[box][m]: box
  // 16 (4 rbh + j) + 4 lk + i (register i), pixel m = 16 (mb0 + c) + li
  zf* tile = reinterpret_cast<zf*>(smem_raw);
  const int n = -(el + eb);
  zd acc[2] = {mkc<double>(0, 0), mkc<double>(0, 0)};
  const bool herm = post.herm != 0;
#pragma unroll 1
  for (int hh = 0; hh < 2; hh++) {
    __syncthreads();
    if (rbh == hh) {
#pragma unroll
      for (int j = 0; j < 4; j++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
          for (int i = 0; i < 4; i++)
            tile[(16 * j + 4 * lk + i) * 64 + 16 * (mb0 + c) + li] =
                mkc<float>(ldexpf(cr[j][c][i], n), ldexpf(ci[j][c][i], n));
    }
    __syncthreads();
    const long o = (long)rhs * rs + (b0 + 64 * hh) * 64;
    if constexpr (POST == 0) {
#pragma unroll 4
      for (int e = tid; e < 64 * 64; e += kBlock) r[o + e] = tile[e];
    } else {
      const zd* gr = post.g + (long)rhs * post.gs + (b0 + 64 * hh) * 64;
      struct In {
        zf x, aux;  // aux: r~ (mode 1)
        zd g;
      };
      stage<2, In>(
          64 * 64,
          [&](int e) {
            In v;
            v.x = post.x[o + e];
            if (POST == 1) v.aux = post.rt[o + e];
            if (herm) v.g = gr[e];  // the forward post step does not use g
            return v;
          },
          [&](int e, In v) {
            const zf gt = tile[e];
            const zd gy = mkc<double>((double)gt.x, (double)gt.y);
            const zd xv = ccast<double>(v.x);
            const zd px = herm ? csub(xv, cconj(cmul(v.g, gy))) : csub(xv, gy);
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
          });
    }
  }
  if constexpr (POST != 0) block_store<2>(acc, post.part + (long)rhs * kSlots * gridDim.x, gridDim.x, blockIdx.x);
}

// P2M of the FP32 operator on split FP16: agm[box][s] = sum_n coeffT[n][s]
// y[box 64 + n] for 32 Morton-consecutive boxes of two right-hand sides.
// The workgroup's input (2 x 32 boxes x 64 pixels) is loaded once into
// registers (units of 8 pixels: 2 per thread), its largest component per
// right-hand side sets the data scale, and it is staged split in LDS as
// [rhs][part][pixel group 8][box 32] x 8 FP16 (boxes XOR 2 (g mod 4), as in
// k_near_h16).  Wave w owns right-hand side w / 2, row block w % 2 (16
// boxes) and all ceil(S / 16) column blocks (samples); K = 64 pixels in two
// MFMA chunks of 32.  The coefficient table coef[kc][part][g][s] (s padded to
// 16 ceil(S / 16), zero beyond S) carries its own scale 2^ec.  Results go
// straight from the accumulators to HBM: lane li stores sample 16 cb + li of
// box 16 j + 4 (lane / 16) + i, 16 consecutive samples per row.
// Grid (B / 32, ceil(nrhs / 2)), kBlock threads.
constexpr int kP2mH16MaxCB = 6;  // column blocks held per wave (S <= 96)
// (at least 4 waves per SIMD: unbounded, the compiler hoists every
// coefficient fragment of both K chunks and takes 256 registers)
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) k_p2m_h16(const zf* __restrict__ y, long ys, zf* __restrict__ agm, long as,
                                                    const v8h* __restrict__ coef, int ec, int S, int nrhs,
                                                    const int* __restrict__ act) {
  __shared__ v8h hal[2 * 4 * 8 * 32];
  __shared__ float red[4][kBlock / 64];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int rw = w >> 1, j = w & 1;
  const int rhs[2] = {2 * (int)blockIdx.y, 2 * (int)blockIdx.y + 1};
  bool on[2];
#pragma unroll
  for (int q = 0; q < 2; q++) on[q] = rhs[q] < nrhs && !rhs_off(act, rhs[q]);
  if (!on[0] && !on[1]) return;
  const bool mine = sel2(on, rw);
  const int rv[2] = {on[0] ? rhs[0] : rhs[1], on[1] ? rhs[1] : rhs[0]};
  const long b0 = (long)xcd_block() * 32;
  const int ncb = (S + 15) / 16, ncolp = 16 * ncb;
  // units (q, box, g8): pixels 8 g8 .. +7 of box b0 + box of right-hand side q; unit t of this thread
  // is tid + 256 t, so q = t and g8 = tid % 8 for both
  const int bx = (tid >> 3) & 31, g8 = tid & 7;
  float4 v[2][4];
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const float4* src = reinterpret_cast<const float4*>(y + (long)sel2(rv, q) * ys + (b0 + bx) * 64 + 8 * g8);
#pragma unroll
    for (int c = 0; c < 4; c++) v[q][c] = src[c];
  }
  int es[2];
  {
    float m[2];
#pragma unroll
    for (int q = 0; q < 2; q++) {
      m[q] = 0.f;
      if (sel2(on, q))
#pragma unroll
        for (int c = 0; c < 4; c++) m[q] = fmaxf(m[q], h16_absmax(v[q][c]));
    }
    block_max<2>(m, red);
    es[0] = h16_scale_exp(m[0]);
    es[1] = h16_scale_exp(m[1]);
  }
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const float s = ldexpf(1.f, sel2(es, q));
    const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    const bool ok = sel2(on, q);
    v8h ph[4];
    h16_split4(ok ? v[q][0] : z, ok ? v[q][1] : z, s, ph, 0);
    h16_split4(ok ? v[q][2] : z, ok ? v[q][3] : z, s, ph, 4);
#pragma unroll
    for (int P = 0; P < 4; P++) hal[((q * 4 + P) * 8 + g8) * 32 + (bx ^ (2 * (g8 & 3)))] = ph[P];
  }
  __syncthreads();
  if (!mine) return;  // no barrier follows
  v4f cr[kP2mH16MaxCB], ci[kP2mH16MaxCB];
#pragma unroll
  for (int c = 0; c < kP2mH16MaxCB; c++) {
    cr[c] = v4f{0, 0, 0, 0};
    ci[c] = v4f{0, 0, 0, 0};
  }
  const v8h* tl = coef + lk * ncolp + li;  // + (kc * 4 + P) * 4 * ncolp + 16 cb
#pragma unroll
  for (int kc = 0; kc < 2; kc++) {
    v8h a[4];
#pragma unroll
    for (int P = 0; P < 4; P++) a[P] = hal[((rw * 4 + P) * 8 + 4 * kc + lk) * 32 + ((16 * j + li) ^ (2 * lk))];
#pragma unroll
    for (int cb = 0; cb < kP2mH16MaxCB; cb++) {
      if (cb < ncb) {
        v8h b[4], nb[2];
#pragma unroll
        for (int P = 0; P < 4; P++) b[P] = tl[(long)(kc * 4 + P) * 4 * ncolp + 16 * cb];
        nb[0] = h16_neg(b[2]);
        nb[1] = h16_neg(b[3]);
        h16_cmac(a, b, nb, cr[cb], ci[cb]);
      }
    }
  }
  const int n = -(sel2(es, rw) + ec);
  zf* out = agm + (long)sel2(rhs, rw) * as + (b0 + 16 * j + 4 * lk) * S;
#pragma unroll
  for (int cb = 0; cb < kP2mH16MaxCB; cb++) {
    const int sm = 16 * cb + li;
    if (cb < ncb && sm < S) {
#pragma unroll
      for (int i = 0; i < 4; i++) out[(long)i * S + sm] = mkc<float>(ldexpf(cr[cb][i], n), ldexpf(ci[cb][i], n));
    }
  }
}

}  // namespace isg
