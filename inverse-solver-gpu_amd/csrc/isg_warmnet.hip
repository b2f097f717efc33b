// isg_warmnet.hip -- kernels and driver of the warm-start network.
#include "isg_warmnet.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>

namespace isg {

namespace {

constexpr int kNB = 256;  // threads per block
constexpr int kCK = 32;   // input channels staged per pass
constexpr int kCKP = kCK + 8;  // padded LDS row: 80 bytes keep the 16-byte loads aligned
constexpr int kCOT = 64;  // output channels per workgroup

typedef short v4s __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (inputs are finite).
__host__ __device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float bf2f(uint16_t b) {
  uint32_t u = (uint32_t)b << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

// Convolution (KS = 1 or 3, stride 1, zero padding KS/2) on the matrix cores.
//   Y[n][p][co] = act(bias[co] + sum_{tap,ci} bf16(In[n][p+tap][ci]) bf16(Wb[tap][co][ci]))
//   In = X (+ up2(U)) (* [M > 0])
// One workgroup: a 16x16-position tile of one image and up to 64 output
// channels (NCB blocks of 16).  Waves split the output blocks and the tile's
// 16 rows: wave w owns block w / (4 / NCB) and rows w % (4 / NCB) + (4 / NCB) j,
// one MFMA row block (16 positions of a tile row) each, so a wave keeps its
// block's weights of one input chunk (KS * KS fragments) in registers and
// reuses them over 16 * NCB / 4 row blocks.  The input halo is staged through
// LDS in chunks of 32 (or 16) input channels as BF16; a chunk of 32 is one
// v_mfma_f32_16x16x32_bf16 per tap and row block (lane l supplies position
// or output channel l % 16 and the 8 input channels 8 (l / 16) .. +7 of both
// operands; the same channel order on both sides, so the sum over the chunk
// is exact in any lane-to-channel assignment), a chunk of 16 one
// v_mfma_f32_16x16x16_bf16.  Positions outside the image are zero on input
// and not stored (partial tiles).
template <int KS, int NCB, int MODE>  // MODE 0: In = X, 1: X + up2(U), 2: X * [M > 0]
__global__ void __launch_bounds__(kNB) __attribute__((amdgpu_waves_per_eu(2))) k_conv(const float* __restrict__ X, const float* __restrict__ U,
                                              const float* __restrict__ M, const uint16_t* __restrict__ Wb,
                                              const float* __restrict__ bias, float* __restrict__ Y, int cin,
                                              int cout, int cstore, int H, int W, int relu) {
  constexpr int TS = 16, R = KS / 2, HW = TS + KS - 1, HS = HW * HW;
  constexpr int WPC = 4 / NCB;     // waves per output block
  constexpr int NRB = TS / WPC;    // row blocks per wave
  __shared__ __attribute__((aligned(16))) uint16_t hal[HS * kCKP];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int n = blockIdx.y;
  const int tiles_x = (W + TS - 1) / TS;  // the last tile row / column may be partial
  const int ty0 = (blockIdx.x / tiles_x) * TS, tx0 = (blockIdx.x % tiles_x) * TS;
  const int cb = w / WPC, part = w % WPC;
  const int co = blockIdx.z * kCOT + cb * 16 + (lane & 15);  // this lane's B column / D channel
  const int ck = (cin % kCK == 0) ? kCK : 16;
  const int li = lane & 15, kq = lane >> 4;
  v4f acc[NRB];
#pragma unroll
  for (int j = 0; j < NRB; j++) acc[j] = v4f{0, 0, 0, 0};
  const long img = (long)n * H * W;
  const uint16_t* wrow = Wb + (long)co * cin;  // + tap * cout * cin + ci
  const int sh4 = ck == kCK ? 3 : 2;  // log2 of the float4 groups per position in a chunk
  for (int ci0 = 0; ci0 < cin; ci0 += ck) {
    // halo staging: every load of the thread's share is issued before the
    // first conversion, so their latencies overlap (NE float4 groups each)
    constexpr int NE = (HS * (kCK / 4) + kNB - 1) / kNB;
    float4 v[NE], a2[MODE ? NE : 1];  // a2: the upsampled addend (MODE 1) or the mask (MODE 2)
#pragma unroll
    for (int q = 0; q < NE; q++) {
      const int e = tid + q * kNB;
      const int hp = e >> sh4, c4 = e & ((1 << sh4) - 1);
      const int gy = ty0 + hp / HW - R, gx = tx0 + hp % HW - R;
      v[q] = make_float4(0.f, 0.f, 0.f, 0.f);
      if constexpr (MODE != 0) a2[q] = v[q];
      if (hp < HS && gy >= 0 && gy < H && gx >= 0 && gx < W) {
        const long pix = img + (long)gy * W + gx;
        v[q] = *reinterpret_cast<const float4*>(X + pix * cin + ci0 + 4 * c4);
        if constexpr (MODE == 1) {
          const long up = (long)n * (H >> 1) * (W >> 1) + (long)(gy >> 1) * (W >> 1) + (gx >> 1);
          a2[q] = *reinterpret_cast<const float4*>(U + up * cin + ci0 + 4 * c4);
        }
        if constexpr (MODE == 2) a2[q] = *reinterpret_cast<const float4*>(M + pix * cin + ci0 + 4 * c4);
      }
    }
#pragma unroll
    for (int q = 0; q < NE; q++) {
      const int e = tid + q * kNB;
      const int hp = e >> sh4, c4 = e & ((1 << sh4) - 1);
      if (hp >= HS) continue;
      float4 x = v[q];
      if constexpr (MODE == 1) {
        x.x += a2[q].x; x.y += a2[q].y; x.z += a2[q].z; x.w += a2[q].w;
      }
      if constexpr (MODE == 2) {
        x.x = a2[q].x > 0.f ? x.x : 0.f;
        x.y = a2[q].y > 0.f ? x.y : 0.f;
        x.z = a2[q].z > 0.f ? x.z : 0.f;
        x.w = a2[q].w > 0.f ? x.w : 0.f;
      }
      uint2 pk;
      pk.x = (uint32_t)f2bf(x.x) | ((uint32_t)f2bf(x.y) << 16);
      pk.y = (uint32_t)f2bf(x.z) | ((uint32_t)f2bf(x.w) << 16);
      *reinterpret_cast<uint2*>(&hal[hp * kCKP + 4 * c4]) = pk;
    }
    // this lane's weight fragment of one tap (8 input channels of its output
    // channel), the next tap's loaded while the current one's MFMAs run
    if (ck == kCK) {
      const uint16_t* wp = wrow + ci0 + 8 * kq;
      v8s bn = *reinterpret_cast<const v8s*>(wp);
      __syncthreads();
#pragma unroll 1
      for (int tap = 0; tap < KS * KS; tap++) {
        const v8bf b = __builtin_bit_cast(v8bf, bn);
        if (tap + 1 < KS * KS) bn = *reinterpret_cast<const v8s*>(wp + (long)(tap + 1) * cout * cin);
        const int dy = tap / KS, dx = tap % KS;
        const uint16_t* hb = &hal[(dy * HW + li + dx) * kCKP + 8 * kq];
#pragma unroll
        for (int j = 0; j < NRB; j++) {
          const int rb = part + WPC * j;
          const v8bf a = __builtin_bit_cast(v8bf, *reinterpret_cast<const v8s*>(hb + rb * HW * kCKP));
          acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[j], 0, 0, 0);
        }
      }
    } else {
      const uint16_t* wp = wrow + ci0 + 4 * kq;
      v4s bn = *reinterpret_cast<const v4s*>(wp);
      __syncthreads();
#pragma unroll 1
      for (int tap = 0; tap < KS * KS; tap++) {
        const v4s b = bn;
        if (tap + 1 < KS * KS) bn = *reinterpret_cast<const v4s*>(wp + (long)(tap + 1) * cout * cin);
        const int dy = tap / KS, dx = tap % KS;
        const uint16_t* hb = &hal[(dy * HW + li + dx) * kCKP + 4 * kq];
#pragma unroll
        for (int j = 0; j < NRB; j++) {
          const int rb = part + WPC * j;
          const v4s a = *reinterpret_cast<const v4s*>(hb + rb * HW * kCKP);
          acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[j], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
  // D[i][j]: lane j + 16 (i / 4), register i % 4; i = position within the row block, j = channel
  const float bv = bias ? bias[co] : 0.f;
#pragma unroll
  for (int j = 0; j < NRB; j++) {
    const int gy = ty0 + part + WPC * j;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int gx = tx0 + 4 * kq + r;
      float v = acc[j][r] + bv;
      if (relu) v = v > 0.f ? v : 0.f;
      if (co < cstore && gy < H && gx < W) Y[(img + (long)gy * W + gx) * cstore + co] = v;
    }
  }
}

// Weight-gradient partial sums of one chunk of pixel tiles:
//   part[chunk][tap][co][ci] = sum_p bf16(dY[p][co] [M>0]) bf16(In[p+tap][ci])
//   bpart[chunk][co]         = sum_p bf16(dY[p][co] [M>0])       (ci tile 0 only)
// with In = X (+ up2(U)).  One workgroup: 16 output x 16 input channels; per
// 8x8 pixel tile the products are v_mfma_f32_16x16x32_bf16 over the tile's
// positions (two k-steps of 32 = four tile rows each): A = dY^T (rows co,
// lane l supplies channel l % 16 and the 8 positions of tile row
// 4 c + l / 16), B = the input shifted by the tap (columns ci, the same
// positions); wave w owns taps w, w + 4, w + 8.  The tile's dY and input halo
// are staged in LDS as BF16, channel-major.  Fixed order over the chunk's
// tiles: deterministic.
template <int KS>
__global__ void __launch_bounds__(kNB) k_wgrad(const float* __restrict__ dY, const float* __restrict__ M,
                                               const float* __restrict__ X, const float* __restrict__ U, int cin,
                                               int cout, int H, int W, int batch, int per_chunk,
                                               float* __restrict__ part, float* __restrict__ bpart) {
  constexpr int R = KS / 2, HW = 8 + KS - 1, HS = HW * HW, NT = KS * KS, TPW = (NT + 3) / 4;
  __shared__ __attribute__((aligned(16))) uint16_t dsm[16][64];  // dY[co][p], BF16
  __shared__ uint16_t xsm[16][HS];                              // In[ci][halo position], BF16
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, li = lane & 15, lk = lane >> 4;
  const int ncit = cin >> 4;
  const int co0 = (blockIdx.x / ncit) * 16, ci0 = (blockIdx.x % ncit) * 16;
  const int chunk = blockIdx.y;
  const int tiles_x = (W + 7) >> 3, tiles = ((H + 7) >> 3) * tiles_x;
  const long total = (long)tiles * batch;
  v4f acc[TPW];
#pragma unroll
  for (int q = 0; q < TPW; q++) acc[q] = v4f{0, 0, 0, 0};
  float bacc = 0.f;
  for (long tt = (long)chunk * per_chunk; tt < total && tt < (long)(chunk + 1) * per_chunk; tt++) {
    const int n = (int)(tt / tiles), tl = (int)(tt % tiles);
    const int ty0 = (tl / tiles_x) * 8, tx0 = (tl % tiles_x) * 8;
    const long img = (long)n * H * W;
    for (int e = tid; e < 64 * 16; e += kNB) {
      const int p = e >> 4, c = e & 15;
      const int gy = ty0 + (p >> 3), gx = tx0 + (p & 7);
      float v = 0.f;
      if (gy < H && gx < W) {
        const long pix = img + (long)gy * W + gx;
        v = dY[pix * cout + co0 + c];
        if (M && !(M[pix * cout + co0 + c] > 0.f)) v = 0.f;
      }
      dsm[c][p] = f2bf(v);
    }
    for (int e = tid; e < HS * 16; e += kNB) {
      const int hp = e >> 4, c = e & 15;
      const int gy = ty0 + hp / HW - R, gx = tx0 + hp % HW - R;
      float v = 0.f;
      if (gy >= 0 && gy < H && gx >= 0 && gx < W) {
        v = X[(img + (long)gy * W + gx) * cin + ci0 + c];
        if (U) v += U[((long)n * (H >> 1) * (W >> 1) + (long)(gy >> 1) * (W >> 1) + (gx >> 1)) * cin + ci0 + c];
      }
      xsm[c][hp] = f2bf(v);
    }
    __syncthreads();
    if (tid < 16)
      for (int p = 0; p < 64; p++) bacc += bf2f(dsm[tid][p]);
#pragma unroll
    for (int q = 0; q < TPW; q++) {
      const int t = w + 4 * q;
      if (t < NT) {
        const int dy = t / KS, dx = t % KS;
#pragma unroll
        for (int c = 0; c < 2; c++) {
          const int row = 4 * c + lk;  // tile row of this lane's 8 positions
          const v8bf a = __builtin_bit_cast(v8bf, *reinterpret_cast<const v8s*>(&dsm[li][8 * row]));
          const uint16_t* xr = &xsm[li][(row + dy) * HW + dx];
          v8s bs;
#pragma unroll
          for (int j = 0; j < 8; j++) bs[j] = (short)xr[j];
          acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, __builtin_bit_cast(v8bf, bs), acc[q], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
  const long base = (long)chunk * NT * cout * cin;
  // D[co][ci]: lane l holds rows co0 + 4 (l / 16) + i, column ci0 + l % 16
#pragma unroll
  for (int q = 0; q < TPW; q++) {
    const int t = w + 4 * q;
    if (t < NT) {
#pragma unroll
      for (int i = 0; i < 4; i++) part[base + ((long)t * cout + co0 + 4 * lk + i) * cin + ci0 + li] = acc[q][i];
    }
  }
  if (ci0 == 0 && tid < 16) bpart[(long)chunk * cout + co0 + tid] = bacc;
}

// out[i] = sum_c part[c][i], fixed order (deterministic)
__global__ void k_chunk_sum(const float* __restrict__ part, long len, int nchunk, float* __restrict__ out) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (long)gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int c = 0; c < nchunk; c++) s += part[(long)c * len + i];
    out[i] = s;
  }
}

// Space-to-depth input: x32[n][Y][X][8 s + c] = in[n][2Y + s / 2][2X + s % 2][c] * scale[c]
// (s = 2 dy + dx, c < 8); H, W are the image's sides.
__global__ void k_net_pack(const float* __restrict__ in, int batch, int H, int W, const float* __restrict__ sc,
                           float* __restrict__ x32) {
  constexpr int C = WarmNet::kInCh, S = WarmNet::kS2D;
  const int Hn = H / S, Wn = W / S;
  const long total = (long)batch * Hn * Wn * S * S * C;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    long q = i / C;
    const int sp = (int)(q % (S * S));
    q /= S * S;
    const int X = (int)(q % Wn);
    q /= Wn;
    const int Y = (int)(q % Hn);
    const long n = q / Hn;
    const long src = ((n * H + (long)S * Y + sp / S) * W + (long)S * X + sp % S) * C + c;
    x32[i] = in[src] * sc[c];
  }
}

// Depth-to-space output: out[n][2Y + s / 2][2X + s % 2][c] = y16[n][Y][X][2 s + c] * scale, c < 2
__global__ void k_net_unpack(const float* __restrict__ y16, int batch, int H, int W, float scale,
                             float* __restrict__ out) {
  constexpr int S = WarmNet::kS2D;
  const int Wn = W / S;
  const long total = (long)batch * H * W * 2;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int c = (int)(i & 1);
    long q = i >> 1;
    const int x = (int)(q % W);
    q /= W;
    const int y = (int)(q % H);
    const long n = q / H;
    const int sp = (y % S) * S + x % S;
    const long pn = (n * (H / S) + y / S) * Wn + x / S;
    out[i] = y16[pn * 16 + 2 * sp + c] * scale;
  }
}

// 2x2 pooling: average (sum = 0) or sum (sum = 1); H, W of the input.
__global__ void k_pool(const float* __restrict__ x, float* __restrict__ y, int batch, int H, int W, int C, int sum) {
  const long n_out = (long)batch * (H / 2) * (W / 2) * C;
  const float f = sum ? 1.f : 0.25f;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n_out; i += (long)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    long p = i / C;
    const int ox = (int)(p % (W / 2));
    p /= (W / 2);
    const int oy = (int)(p % (H / 2));
    const long n = p / (H / 2);
    const long b = ((n * H + 2 * oy) * W + 2 * ox) * C + c;
    y[i] = f * (x[b] + x[b + C] + x[b + (long)W * C] + x[b + (long)W * C + C]);
  }
}

// g[n][y][x][c] += 0.25 gp[n][y/2][x/2][c]  (average-pool backward)
__global__ void k_unpool_add(float* __restrict__ g, const float* __restrict__ gp, int batch, int H, int W, int C) {
  const long n_all = (long)batch * H * W * C;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n_all; i += (long)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    long p = i / C;
    const int x = (int)(p % W);
    p /= W;
    const int y = (int)(p % H);
    const long n = p / H;
    g[i] += 0.25f * gp[((n * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c];
  }
}

// Mean squared error of output channels 0..7 of y16 (2 per sub-pixel) against
// target * inv (target [n][H][W][2], the image's sides), gradient into g16
// (zero on channels 8..15) and per-block partial sums of squares.
__global__ void __launch_bounds__(kNB) k_loss(const float* __restrict__ y16, const float* __restrict__ t, int batch,
                                              int H, int W, float inv, float* __restrict__ g16,
                                              double* __restrict__ part) {
  constexpr int S = WarmNet::kS2D;
  const int Hn = H / S, Wn = W / S;
  const long npos = (long)batch * Hn * Wn;
  const double scale = 2.0 / (double)((long)batch * H * W * 2);
  double s = 0;
  for (long i = (long)blockIdx.x * kNB + threadIdx.x; i < npos * 16; i += (long)gridDim.x * kNB) {
    const long p = i >> 4;
    const int j = (int)(i & 15);
    float gv = 0.f;
    if (j < 2 * S * S) {
      const int sp = j >> 1, c = j & 1;
      const int X = (int)(p % Wn);
      const long q = p / Wn;
      const int Y = (int)(q % Hn);
      const long n = q / Hn;
      const long ti = ((n * H + (long)S * Y + sp / S) * W + (long)S * X + sp % S) * 2 + c;
      const float d = y16[i] - t[ti] * inv;
      s += (double)d * d;
      gv = (float)(scale * d);
    }
    g16[i] = gv;
  }
  __shared__ double sh[kNB / 64];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0;
    for (int j = 0; j < kNB / 64; j++) a += sh[j];
    part[blockIdx.x] = a;
  }
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, long n, float lr, float c1, float c2) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float gv = g[i];
    const float a = 0.9f * m1[i] + 0.1f * gv;
    const float b = 0.999f * m2[i] + 0.001f * gv * gv;
    m1[i] = a;
    m2[i] = b;
    w[i] -= lr * (a / c1) / (sqrtf(b / c2) + 1e-8f);
  }
}

// BF16 copies of one layer's weights: forward [tap][co][ci] and the
// input-gradient layout [tap'][ci][co] with tap' the point-reflected tap.
__global__ void k_refresh(const float* __restrict__ w, uint16_t* __restrict__ wb, uint16_t* __restrict__ wt, int ks,
                          int cout, int cin) {
  const long n = (long)ks * ks * cout * cin;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int ci = (int)(i % cin);
    const long r = i / cin;
    const int co = (int)(r % cout), tap = (int)(r / cout);
    const uint16_t b = f2bf(w[i]);
    wb[i] = b;
    const int tf = ks * ks - 1 - tap;
    wt[((long)tf * cin + ci) * cout + co] = b;
  }
}

int grid_for(long n) { return (int)std::min<long>(4096, std::max<long>(1, (n + kNB - 1) / kNB)); }

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace

#define WN_HIP(call)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess && err_.empty()) err_ = std::string("HIP: ") + hipGetErrorString(e_) + " at " #call; \
  } while (0)

const std::vector<NetLayer>& WarmNet::layers() {
  static const std::vector<NetLayer> L = [] {
    // ks, cin, cout, level, relu
    const int t[kLayers][5] = {{3, 32, 32, 0, 1},   {3, 32, 64, 1, 1},   {3, 64, 128, 2, 1}, {3, 128, 256, 3, 1},
                               {1, 256, 128, 3, 0}, {3, 128, 128, 2, 1}, {1, 128, 64, 2, 0}, {3, 64, 64, 1, 1},
                               {1, 64, 32, 1, 0},   {3, 32, 32, 0, 1},   {1, 32, 16, 0, 0}};
    std::vector<NetLayer> v;
    size_t off = 0;
    for (auto& r : t) {
      NetLayer l{r[0], r[1], r[2], r[3], r[4], 0, 0};
      l.woff = off;
      off += (size_t)l.ks * l.ks * l.cout * l.cin;
      l.boff = off;
      off += l.cout;
      v.push_back(l);
    }
    return v;
  }();
  return L;
}

size_t WarmNet::num_params() {
  const NetLayer& l = layers().back();
  return l.boff + l.cout;
}

void* WarmNet::dalloc(size_t bytes) {
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes ? bytes : 16);
  if (e != hipSuccess) {
    if (err_.empty()) err_ = std::string("hipMalloc: ") + hipGetErrorString(e);
    return nullptr;
  }
  allocs_.push_back(p);
  return p;
}

WarmNet::WarmNet(int max_batch, long max_pixels, uint64_t seed) : max_batch_(max_batch), max_pix_(max_pixels) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    err_ = "no HIP device available";
    return;
  }
  if (max_batch < 1 || max_pixels < (long)kTile * kTile || max_pixels % ((long)kTile * kTile) != 0) {
    err_ = "warm-start network: max_batch >= 1 and max_pixels a multiple of 16x16 required";
    return;
  }
  const size_t np = num_params();
  w_ = (float*)dalloc(np * 4);
  g_ = (float*)dalloc(np * 4);
  m1_ = (float*)dalloc(np * 4);
  m2_ = (float*)dalloc(np * 4);
  wb_ = (uint16_t*)dalloc(np * 2);
  wt_ = (uint16_t*)dalloc(np * 2);
  // activations: channels and level of each buffer
  // (levels of the network, whose level 0 is the space-to-depth image)
  const int ach[A_N][2] = {{32, 0},  {32, 0},  {32, 1},  {64, 1},  {64, 2}, {128, 2}, {128, 3}, {256, 3},
                           {128, 3}, {128, 2}, {64, 2}, {64, 1}, {32, 1}, {32, 0},  {16, 0}};
  const int gch[G_N][2] = {{16, 0},  {32, 0}, {32, 0}, {32, 1},  {64, 1}, {64, 1}, {64, 2},
                           {128, 2}, {128, 2}, {128, 3}, {256, 3}, {128, 3}, {64, 2}, {32, 1}};
  const long pix = (long)max_batch * max_pixels / (kS2D * kS2D);
  for (int a = 0; a < A_N; a++) act_[a] = (float*)dalloc(sizeof(float) * ach[a][0] * (pix >> (2 * ach[a][1])));
  for (int a = 0; a < G_N; a++) grd_[a] = (float*)dalloc(sizeof(float) * gch[a][0] * (pix >> (2 * gch[a][1])));
  // weight-gradient partials: at most 1024 workgroups' worth per layer
  part_len_ = 0;
  for (auto& l : layers()) {
    const long pairs = (long)(l.cout / 16) * (l.cin / 16);
    const long nch = std::max<long>(1, 1024 / pairs);
    part_len_ = std::max<size_t>(part_len_, (size_t)nch * (l.ks * l.ks * l.cout * l.cin + l.cout));
  }
  part_ = (float*)dalloc(part_len_ * 4);
  nlpart_ = 1024;
  lpart_ = (double*)dalloc(sizeof(double) * nlpart_);
  dscale_ = (float*)dalloc(sizeof(float) * 16);
  WN_HIP(hipHostMalloc((void**)&lpart_h_, sizeof(double) * nlpart_, hipHostMallocDefault));
  if (!ok()) return;
  // He-normal initialisation (seeded); the unused output channels of L10
  // start (and stay) zero; the output layer starts small so the first
  // correction is small.
  std::vector<float> w(np, 0.f);
  uint64_t s = seed;
  for (int li = 0; li < kLayers; li++) {
    const NetLayer& l = layers()[li];
    const int cin_used = l.cin;
    const int cout_used = li == kLayers - 1 ? kOutCh * kS2D * kS2D : l.cout;
    const double sd = std::sqrt(2.0 / (l.ks * l.ks * cin_used)) * (li == kLayers - 1 ? 0.1 : 1.0);
    for (int tap = 0; tap < l.ks * l.ks; tap++)
      for (int co = 0; co < cout_used; co++)
        for (int ci = 0; ci < cin_used; ci++) {
          double u1 = ((splitmix(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
          double u2 = ((splitmix(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
          double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(2 * M_PI * u2);
          w[l.woff + ((size_t)tap * l.cout + co) * l.cin + ci] = (float)(sd * z);
        }
  }
  set_params(w);
}

WarmNet::~WarmNet() {
  (void)hipDeviceSynchronize();
  for (void* p : allocs_) (void)hipFree(p);
  if (lpart_h_) (void)hipHostFree(lpart_h_);
}

void WarmNet::refresh_bf16(hipStream_t s) {
  for (auto& l : layers()) {
    const long n = (long)l.ks * l.ks * l.cout * l.cin;
    k_refresh<<<grid_for(n), kNB, 0, s>>>(w_ + l.woff, wb_ + l.woff, wt_ + l.woff, l.ks, l.cout, l.cin);
  }
  WN_HIP(hipGetLastError());
}

void WarmNet::get_params(std::vector<float>& out) {
  out.resize(num_params());
  WN_HIP(hipDeviceSynchronize());
  WN_HIP(hipMemcpy(out.data(), w_, out.size() * 4, hipMemcpyDeviceToHost));
}

void WarmNet::get_grads(std::vector<float>& out) {
  out.resize(num_params());
  WN_HIP(hipDeviceSynchronize());
  WN_HIP(hipMemcpy(out.data(), g_, out.size() * 4, hipMemcpyDeviceToHost));
}

void WarmNet::set_params(const std::vector<float>& in) {
  if (in.size() != num_params()) {
    if (err_.empty()) err_ = "warm-start network: parameter count mismatch";
    return;
  }
  WN_HIP(hipDeviceSynchronize());
  WN_HIP(hipMemcpy(w_, in.data(), in.size() * 4, hipMemcpyHostToDevice));
  WN_HIP(hipMemset(g_, 0, in.size() * 4));
  WN_HIP(hipMemset(m1_, 0, in.size() * 4));
  WN_HIP(hipMemset(m2_, 0, in.size() * 4));
  step_ = 0;
  refresh_bf16(nullptr);
  WN_HIP(hipDeviceSynchronize());
}

void WarmNet::conv(int li, const float* x, const float* u, const float* m, float* y, int cstore, bool transpose,
                   int batch, int H, int W, hipStream_t s) {
  const NetLayer& l = layers()[li];
  const int cin = transpose ? l.cout : l.cin, cout = transpose ? l.cin : l.cout;
  const uint16_t* wb = (transpose ? wt_ : wb_) + l.woff;
  const float* bias = transpose ? nullptr : w_ + l.boff;
  const int relu = transpose ? 0 : l.relu;
  dim3 grid(((H + 15) / 16) * ((W + 15) / 16), batch, (cout + kCOT - 1) / kCOT);
  // output blocks of 16 channels per workgroup: 1, 2 or 4 (every layer's widths are 16, 32 or multiples of 64)
  const int ncb = std::min(4, cout / 16);
  // input mode: plain, plus the upsampled decoder input (skip connection), or masked by a ReLU output
  if (u && m) {
    if (err_.empty()) err_ = "warm-start network: a convolution input cannot be both skip-added and masked";
    return;
  }
  const int mode = u ? 1 : (m ? 2 : 0);
#define WN_CONV(KS, NCB, MODE) \
  k_conv<KS, NCB, MODE><<<grid, kNB, 0, s>>>(x, u, m, wb, bias, y, cin, cout, cstore, H, W, relu)
#define WN_CONV_NCB(KS, MODE) \
  do { if (ncb == 4) WN_CONV(KS, 4, MODE); else if (ncb == 2) WN_CONV(KS, 2, MODE); else WN_CONV(KS, 1, MODE); } while (0)
#define WN_CONV_MODE(KS) \
  do { if (mode == 1) WN_CONV_NCB(KS, 1); else if (mode == 2) WN_CONV_NCB(KS, 2); else WN_CONV_NCB(KS, 0); } while (0)
  if (l.ks == 3)
    WN_CONV_MODE(3);
  else
    WN_CONV_MODE(1);
#undef WN_CONV_MODE
#undef WN_CONV_NCB
#undef WN_CONV
}

void WarmNet::wgrad(int li, const float* dy, const float* m, const float* x, const float* u, int batch, int H,
                    int W, hipStream_t s) {
  const NetLayer& l = layers()[li];
  const long pairs = (long)(l.cout / 16) * (l.cin / 16);
  const long tiles = (long)batch * ((H + 7) / 8) * ((W + 7) / 8);
  const long nch = std::min<long>(tiles, std::max<long>(1, 1024 / pairs));
  const int per = (int)((tiles + nch - 1) / nch);
  const long wlen = (long)l.ks * l.ks * l.cout * l.cin;
  float* bpart = part_ + nch * wlen;
  dim3 grid((unsigned)pairs, (unsigned)nch);
  if (l.ks == 3)
    k_wgrad<3><<<grid, kNB, 0, s>>>(dy, m, x, u, l.cin, l.cout, H, W, batch, per, part_, bpart);
  else
    k_wgrad<1><<<grid, kNB, 0, s>>>(dy, m, x, u, l.cin, l.cout, H, W, batch, per, part_, bpart);
  k_chunk_sum<<<grid_for(wlen), kNB, 0, s>>>(part_, wlen, (int)nch, g_ + l.woff);
  k_chunk_sum<<<grid_for(l.cout), kNB, 0, s>>>(bpart, l.cout, (int)nch, g_ + l.boff);
}

static bool shape_ok(int batch, int H, int W, int max_batch, long max_pix) {
  constexpr int T = WarmNet::kTile;
  return batch >= 1 && batch <= max_batch && H >= T && W >= T && H % T == 0 && W % T == 0 && (long)H * W <= max_pix;
}

static const char* kShapeMsg = "warm-start network: image shape outside the workspace (H, W multiples of 16)";

// H, W: the image's sides; the network runs on the space-to-depth image.
void WarmNet::forward_all(const float* in, int batch, int Hi, int Wi, hipStream_t s) {
  WN_HIP(hipMemcpyAsync(dscale_, scales_, sizeof(float) * kScales, hipMemcpyHostToDevice, s));
  const int H = Hi / kS2D, W = Wi / kS2D;
  const long np = (long)batch * H * W;
  k_net_pack<<<grid_for(np * 32), kNB, 0, s>>>(in, batch, Hi, Wi, dscale_, act_[A_X]);
  forward_layers(batch, H, W, s);
}

float* WarmNet::packed_input(int batch, int Hi, int Wi) {
  if (!ok()) return nullptr;
  if (!shape_ok(batch, Hi, Wi, max_batch_, max_pix_)) {
    err_ = kShapeMsg;
    return nullptr;
  }
  return act_[A_X];
}

void WarmNet::forward_packed(int batch, int Hi, int Wi, hipStream_t s) {
  if (!ok()) return;
  if (!shape_ok(batch, Hi, Wi, max_batch_, max_pix_)) {
    err_ = kShapeMsg;
    return;
  }
  forward_layers(batch, Hi / kS2D, Wi / kS2D, s);
  WN_HIP(hipGetLastError());
}

// every layer from the packed input act_[A_X] (H, W: the network's level-0 sides)
void WarmNet::forward_layers(int batch, int H, int W, hipStream_t s) {
  const long np = (long)batch * H * W;
  const int H1 = H / 2, W1 = W / 2, H2 = H / 4, W2 = W / 4, H3 = H / 8, W3 = W / 8;
  conv(0, act_[A_X], nullptr, nullptr, act_[A_E0], 32, false, batch, H, W, s);
  k_pool<<<grid_for(np * 32 / 4), kNB, 0, s>>>(act_[A_E0], act_[A_P0], batch, H, W, 32, 0);
  conv(1, act_[A_P0], nullptr, nullptr, act_[A_E1], 64, false, batch, H1, W1, s);
  k_pool<<<grid_for(np * 64 / 16), kNB, 0, s>>>(act_[A_E1], act_[A_P1], batch, H1, W1, 64, 0);
  conv(2, act_[A_P1], nullptr, nullptr, act_[A_E2], 128, false, batch, H2, W2, s);
  k_pool<<<grid_for(np * 128 / 64), kNB, 0, s>>>(act_[A_E2], act_[A_P2], batch, H2, W2, 128, 0);
  conv(3, act_[A_P2], nullptr, nullptr, act_[A_E3], 256, false, batch, H3, W3, s);
  conv(4, act_[A_E3], nullptr, nullptr, act_[A_U3], 128, false, batch, H3, W3, s);
  conv(5, act_[A_E2], act_[A_U3], nullptr, act_[A_D2], 128, false, batch, H2, W2, s);
  conv(6, act_[A_D2], nullptr, nullptr, act_[A_U2], 64, false, batch, H2, W2, s);
  conv(7, act_[A_E1], act_[A_U2], nullptr, act_[A_D1], 64, false, batch, H1, W1, s);
  conv(8, act_[A_D1], nullptr, nullptr, act_[A_U1], 32, false, batch, H1, W1, s);
  conv(9, act_[A_E0], act_[A_U1], nullptr, act_[A_D0], 32, false, batch, H, W, s);
  conv(10, act_[A_D0], nullptr, nullptr, act_[A_OUT], 16, false, batch, H, W, s);
  WN_HIP(hipGetLastError());
}


void WarmNet::forward(const float* in, float* out, int batch, int H, int W, hipStream_t s) {
  if (!ok()) return;
  if (!shape_ok(batch, H, W, max_batch_, max_pix_)) {
    err_ = kShapeMsg;
    return;
  }
  forward_all(in, batch, H, W, s);
  k_net_unpack<<<grid_for((long)batch * H * W * 2), kNB, 0, s>>>(act_[A_OUT], batch, H, W, scales_[kInCh], out);
  WN_HIP(hipGetLastError());
}

double WarmNet::loss_grad(const float* in, const float* target, int batch, int Hi, int Wi, hipStream_t s) {
  if (!ok()) return 0;
  if (!shape_ok(batch, Hi, Wi, max_batch_, max_pix_)) {
    err_ = kShapeMsg;
    return 0;
  }
  forward_all(in, batch, Hi, Wi, s);
  const int H = Hi / kS2D, W = Wi / kS2D;
  const long np = (long)batch * H * W;
  const int H1 = H / 2, W1 = W / 2, H2 = H / 4, W2 = W / 4, H3 = H / 8, W3 = W / 8;
  const int nb = std::min(nlpart_, grid_for(np * 16));
  k_loss<<<nb, kNB, 0, s>>>(act_[A_OUT], target, batch, Hi, Wi, 1.0f / scales_[kInCh], grd_[G_OUT], lpart_);
  WN_HIP(hipMemcpyAsync(lpart_h_, lpart_, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
  // decoder, top-down
  wgrad(10, grd_[G_OUT], nullptr, act_[A_D0], nullptr, batch, H, W, s);
  conv(10, grd_[G_OUT], nullptr, nullptr, grd_[G_D0], 32, true, batch, H, W, s);
  wgrad(9, grd_[G_D0], act_[A_D0], act_[A_E0], act_[A_U1], batch, H, W, s);
  conv(9, grd_[G_D0], nullptr, act_[A_D0], grd_[G_IN0], 32, true, batch, H, W, s);
  k_pool<<<grid_for(np * 32 / 4), kNB, 0, s>>>(grd_[G_IN0], grd_[G_U1], batch, H, W, 32, 1);
  wgrad(8, grd_[G_U1], nullptr, act_[A_D1], nullptr, batch, H1, W1, s);
  conv(8, grd_[G_U1], nullptr, nullptr, grd_[G_D1], 64, true, batch, H1, W1, s);
  wgrad(7, grd_[G_D1], act_[A_D1], act_[A_E1], act_[A_U2], batch, H1, W1, s);
  conv(7, grd_[G_D1], nullptr, act_[A_D1], grd_[G_IN1], 64, true, batch, H1, W1, s);
  k_pool<<<grid_for(np * 64 / 16), kNB, 0, s>>>(grd_[G_IN1], grd_[G_U2], batch, H1, W1, 64, 1);
  wgrad(6, grd_[G_U2], nullptr, act_[A_D2], nullptr, batch, H2, W2, s);
  conv(6, grd_[G_U2], nullptr, nullptr, grd_[G_D2], 128, true, batch, H2, W2, s);
  wgrad(5, grd_[G_D2], act_[A_D2], act_[A_E2], act_[A_U3], batch, H2, W2, s);
  conv(5, grd_[G_D2], nullptr, act_[A_D2], grd_[G_IN2], 128, true, batch, H2, W2, s);
  k_pool<<<grid_for(np * 128 / 64), kNB, 0, s>>>(grd_[G_IN2], grd_[G_U3], batch, H2, W2, 128, 1);
  wgrad(4, grd_[G_U3], nullptr, act_[A_E3], nullptr, batch, H3, W3, s);
  conv(4, grd_[G_U3], nullptr, nullptr, grd_[G_E3], 256, true, batch, H3, W3, s);
  // encoder, bottom-up
  wgrad(3, grd_[G_E3], act_[A_E3], act_[A_P2], nullptr, batch, H3, W3, s);
  conv(3, grd_[G_E3], nullptr, act_[A_E3], grd_[G_P2], 128, true, batch, H3, W3, s);
  k_unpool_add<<<grid_for(np * 128 / 16), kNB, 0, s>>>(grd_[G_IN2], grd_[G_P2], batch, H2, W2, 128);
  wgrad(2, grd_[G_IN2], act_[A_E2], act_[A_P1], nullptr, batch, H2, W2, s);
  conv(2, grd_[G_IN2], nullptr, act_[A_E2], grd_[G_P1], 64, true, batch, H2, W2, s);
  k_unpool_add<<<grid_for(np * 64 / 4), kNB, 0, s>>>(grd_[G_IN1], grd_[G_P1], batch, H1, W1, 64);
  wgrad(1, grd_[G_IN1], act_[A_E1], act_[A_P0], nullptr, batch, H1, W1, s);
  conv(1, grd_[G_IN1], nullptr, act_[A_E1], grd_[G_P0], 32, true, batch, H1, W1, s);
  k_unpool_add<<<grid_for(np * 32), kNB, 0, s>>>(grd_[G_IN0], grd_[G_P0], batch, H, W, 32);
  wgrad(0, grd_[G_IN0], act_[A_E0], act_[A_X], nullptr, batch, H, W, s);
  WN_HIP(hipGetLastError());
  WN_HIP(hipStreamSynchronize(s));
  double loss = 0;
  for (int j = 0; j < nb; j++) loss += lpart_h_[j];
  return loss / (double)((long)batch * Hi * Wi * 2);
}

void WarmNet::adam(float lr, hipStream_t s) {
  if (!ok()) return;
  step_++;
  const float c1 = 1.f - std::pow(0.9f, (float)step_), c2 = 1.f - std::pow(0.999f, (float)step_);
  const long n = (long)num_params();
  k_adam<<<grid_for(n), kNB, 0, s>>>(w_, g_, m1_, m2_, n, lr, c1, c2);
  refresh_bf16(s);
}

// File: "ISGWNET2", uint32 parameter count, 9 float scales, float parameters.
bool WarmNet::save(const char* path) {
  std::vector<float> w;
  get_params(w);
  if (!ok()) return false;
  FILE* f = std::fopen(path, "wb");
  if (!f) {
    err_ = std::string("cannot write ") + path;
    return false;
  }
  uint32_t n = (uint32_t)w.size();
  bool good = std::fwrite("ISGWNET2", 1, 8, f) == 8 && std::fwrite(&n, 4, 1, f) == 1 &&
              std::fwrite(scales_, 4, kScales, f) == (size_t)kScales && std::fwrite(w.data(), 4, n, f) == n;
  good = (std::fclose(f) == 0) && good;
  if (!good) err_ = std::string("write failed: ") + path;
  return good;
}

bool WarmNet::load(const char* path) {
  FILE* f = std::fopen(path, "rb");
  if (!f) {
    err_ = std::string("cannot read ") + path;
    return false;
  }
  char magic[8];
  uint32_t n = 0;
  float sc[kScales];
  std::vector<float> w;
  bool good = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "ISGWNET2", 8) == 0 &&
              std::fread(&n, 4, 1, f) == 1 && n == num_params() && std::fread(sc, 4, kScales, f) == (size_t)kScales;
  if (good) {
    w.resize(n);
    good = std::fread(w.data(), 4, n, f) == n;
  }
  std::fclose(f);
  if (!good) {
    err_ = std::string("not a warm-start weight file for this network: ") + path;
    return false;
  }
  set_scales(sc);
  set_params(w);
  return ok();
}

}  // namespace isg
