// isg_engine.hip -- device tables, MLFMA orchestration, receivers and the
// batched BiCGSTAB driver of the MI355X solver.
#include "isg_engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <type_traits>

#include <rocblas/rocblas.h>

#include "isg_kernels.hpp"
#include "isg_near_h16.hpp"
#include "isg_setup_kernels.hpp"
#include "isg_solver_kernels.hpp"

namespace isg {

#define ISG_HIP(call)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess) {                                                                \
      set_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " #call);        \
    }                                                                                      \
  } while (0)

// Device scratch of the allocation-free solve (Engine::bicgs_static,
// isg_bicgs_dev_static): the solver's reduction partials, per-system scalars,
// activity flags, tolerances and blend coefficients.  One copy per GPU, held
// until the process ends; host code reaches it only through the address
// hipGetSymbolAddress returns for the current device (g_scratch_ptr).
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {
inline unsigned nblocks(long work) { return (unsigned)((work + kBlock - 1) / kBlock); }

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// bytes[0..n) = 0 (a fill kernel, not hipMemsetAsync, so that a caller may capture it)
__global__ void __launch_bounds__(kBlock) k_clear_words(unsigned* __restrict__ p, long n) {
  for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock) p[i] = 0u;
}

// Finite-radius receivers (line-source geometry): K[m][n] = (i/4) H0^(1)(k0
// |rho_m - r_n|) res^2 with rho_m = radius (cos, sin)(2 pi m / R), the
// midpoint-rule Green's function weight of pixel n seen from receiver m,
// stored in chunks of kc unknowns: Kb[b][m][j] = K[m][b kc + j].  (rocBLAS's
// GEMM kernels slow down threefold once one operand spans 4 GiB -- the
// unchunked 2048^2 x 512 operator is 34 GB -- so every operand is kept
// chunk-compact.)
__global__ void __launch_bounds__(kBlock) k_rx_matrix(const zd* __restrict__ pos, long N, int R, long kc, double radius,
                                                      double k0, double res2, zd* __restrict__ Kb) {
  const int m = blockIdx.y;
  double sa, ca;
  sincos(2.0 * M_PI * m / R, &sa, &ca);
  const double rx = radius * ca, ry = radius * sa;
  for (long n = (long)blockIdx.x * kBlock + threadIdx.x; n < N; n += (long)gridDim.x * kBlock) {
    const double dx = rx - pos[n].x, dy = ry - pos[n].y;
    const double kd = k0 * sqrt(dx * dx + dy * dy);
    const long b = n / kc, j = n - b * kc;
    // (i/4)(J0 + i Y0) res^2
    Kb[(b * R + m) * kc + j] = mkc<double>(-0.25 * y0(kd) * res2, 0.25 * j0(kd) * res2);
  }
}

// xb[b][t][j] = x[t][b kc + j], t < T (the receiver product's input in the operator's chunking)
__global__ void __launch_bounds__(kBlock) k_rx_chunk(const zd* __restrict__ x, long N, int T, long kc,
                                                     zd* __restrict__ xb) {
  const long tot = N * T;
  for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < tot; i += (long)gridDim.x * kBlock) {
    const long t = i / N, n = i - t * N, b = n / kc, j = n - b * kc;
    xb[(b * T + t) * kc + j] = x[i];
  }
}

// out[i] = sum_b part[b][i] (fixed order) -- split-K reduction of the receiver product
__global__ void __launch_bounds__(kBlock) k_sum_parts(const zd* __restrict__ part, long len, int nb, zd* __restrict__ out) {
  for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < len; i += (long)gridDim.x * kBlock) {
    zd s = mkc<double>(0, 0);
    for (int b = 0; b < nb; b++) s = cadd(s, part[(long)b * len + i]);
    out[i] = s;
  }
}

// out[i] = conj(in[i]); in and out may be the same array
__global__ void __launch_bounds__(kBlock) k_conj_copy(const zd* in, zd* out, long len) {
  for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < len; i += (long)gridDim.x * kBlock) out[i] = cconj(in[i]);
}

enum { V_R = 0, V_RT, V_P, V_V, V_S, V_T, V_Y, V_OP, V_RES, V_COR, V_COUNT };
// largest power of two <= v (v >= 1)
int pow2_floor(int v) {
  int p = 1;
  while (p * 2 <= v) p <<= 1;
  return p;
}
constexpr size_t kLdsLimit = 64 * 1024;
constexpr int kQM2M = 12, kQL2L = 24;  // band windows up to 48 (M2M) / 96 (L2L) samples

// Band blocks of a banded M x K matrix whose row k has taps t < nt on input
// (first[k] + t) mod K with weight wt[t * M + k]: for each block of 16 rows,
// the window start s0 and the dense 16 x KB block AB[rb][kk][16].  Returns KB
// (multiple of 4) or 0 when some block's taps do not fit in kbmax <= K inputs.
int build_band(int M, int K, int nt, const std::vector<int>& first, const std::vector<double>& wt, int kbmax,
               std::vector<double>& AB, std::vector<int>& s0) {
  const int nrb = (M + 15) / 16;
  s0.assign(nrb, 0);
  int need = 0;
  auto md = [K](long v) { return (int)(((v % K) + K) % K); };
  for (int rb = 0; rb < nrb; rb++) {
    s0[rb] = md(first[16 * rb]);
    for (int i = 0; i < 16 && 16 * rb + i < M; i++) need = std::max(need, md(first[16 * rb + i] - s0[rb]) + nt);
  }
  const int KB = (need + 3) / 4 * 4;
  if (KB > std::min(K, kbmax)) return 0;
  AB.assign((size_t)nrb * KB * 16, 0.0);
  for (int rb = 0; rb < nrb; rb++)
    for (int i = 0; i < 16 && 16 * rb + i < M; i++) {
      const int k = 16 * rb + i, off = md(first[k] - s0[rb]);
      for (int t = 0; t < nt; t++) AB[((size_t)rb * KB + off + t) * 16 + i] += wt[(size_t)t * M + k];
    }
  return KB;
}
constexpr int kRefineRounds = 3;

// Largest LDS window (source samples) of a target chunk for the chunked
// interpolation kernels, or -1 when the first-tap indices of some chunk are
// not monotone (mod the source count) and the window formula would not hold.
int chunk_window(const std::vector<int>& first, int ntgt, int nsrc, int ntap, int kch) {
  int wmax = 0;
  for (int k0 = 0; k0 < ntgt; k0 += kch) {
    int kn = std::min(kch, ntgt - k0), prev = 0;
    for (int k = k0; k < k0 + kn; k++) {
      int d = first[k] - first[k0];
      if (d < 0) d += nsrc;
      if (d < prev) return -1;
      prev = d;
    }
    wmax = std::max(wmax, prev + ntap);
  }
  return wmax;
}
}  // namespace

template <>
DevOp<double>& Engine::ops<double>() {
  return op64_;
}
template <>
DevOp<float>& Engine::ops<float>() {
  return op32_;
}

void* Engine::dalloc(size_t bytes) {
  void* p = nullptr;
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc failed: ") + hipGetErrorString(e));
    return nullptr;
  }
  allocs_.push_back(p);
  return p;
}

template <typename T>
void Engine::upload(DevOp<T>& d) {
  auto up_c = [&](const std::vector<cd>& src) -> cplx<T>* {
    std::vector<cplx<T>> tmp(src.size());
    for (size_t i = 0; i < src.size(); i++) tmp[i] = mkc<T>((T)src[i].real(), (T)src[i].imag());
    cplx<T>* p = (cplx<T>*)dalloc(tmp.size() * sizeof(cplx<T>));
    if (p) ISG_HIP(hipMemcpy(p, tmp.data(), tmp.size() * sizeof(cplx<T>), hipMemcpyHostToDevice));
    return p;
  };
  auto up_r = [&](const std::vector<double>& src) -> T* {
    std::vector<T> tmp(src.begin(), src.end());
    T* p = (T*)dalloc(tmp.size() * sizeof(T));
    if (p) ISG_HIP(hipMemcpy(p, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice));
    return p;
  };
  int L = P_.L;
  d.interpT.assign(L, nullptr);
  d.anterpT.assign(L, nullptr);
  d.shiftM.assign(L, nullptr);
  d.shiftL.assign(L, nullptr);
  d.trans.assign(L, nullptr);
  if (cfg_.plan.host_tables) {
    d.near = up_c(P_.near);
    d.coeffT = up_c(P_.coeffT);
    d.basisT = up_c(P_.basisT);
    for (int i = 0; i < L; i++) {
      if (!P_.shiftM[i].empty()) d.shiftM[i] = up_c(P_.shiftM[i]);
      if (!P_.shiftL[i].empty()) d.shiftL[i] = up_c(P_.shiftL[i]);
      if (!P_.trans[i].empty()) d.trans[i] = up_c(P_.trans[i]);
    }
  } else if constexpr (sizeof(T) == 8) {
    device_tables(d);
  } else {
    // the FP32 operator's tables: the FP64 device tables rounded
    auto cvt = [&](const zd* src, size_t n) -> cplx<T>* {
      cplx<T>* p = (cplx<T>*)dalloc(n * sizeof(cplx<T>));
      if (p && src)
        setup::k_table_f32<<<(unsigned)std::min<size_t>(4096, (n + 255) / 256), 256, 0, stream_>>>(src, p, (long)n);
      return p;
    };
    const int Sl = P_.nsamp[L - 1];
    d.near = cvt(op64_.near, (size_t)9 * 64 * 64);
    d.coeffT = cvt(op64_.coeffT, (size_t)64 * Sl);
    d.basisT = cvt(op64_.basisT, (size_t)Sl * 64);
    for (int i = 0; i < L; i++) {
      if (op64_.shiftM[i]) d.shiftM[i] = cvt(op64_.shiftM[i], (size_t)4 * P_.nsamp[i]);
      if (op64_.shiftL[i]) d.shiftL[i] = cvt(op64_.shiftL[i], (size_t)4 * P_.nsamp[i]);
      if (op64_.trans[i]) d.trans[i] = cvt(op64_.trans[i], (size_t)49 * P_.nsamp[i]);
    }
    ISG_HIP(hipGetLastError());
    ISG_HIP(hipStreamSynchronize(stream_));
  }
  for (int i = 0; i < L; i++) {
    if (!P_.interpT[i].empty()) d.interpT[i] = up_r(P_.interpT[i]);
    if (!P_.anterpT[i].empty()) d.anterpT[i] = up_r(P_.anterpT[i]);
  }
  d.rxT = up_r(P_.rxT);
  d.rxaT = up_r(P_.rxaT);
  size_t tot = lvl_off_[L - 1] + lvl_size_[L - 1] * cfg_.max_rhs;
  d.agm = (cplx<T>*)dalloc(tot * sizeof(cplx<T>));
  d.agl = (cplx<T>*)dalloc(tot * sizeof(cplx<T>));
}

Engine::Engine(const EngineConfig& cfg) : cfg_(cfg) {
  if (cfg_.max_rhs < 1) cfg_.max_rhs = 1;
  if (!build_plan(cfg_.plan, P_)) {
    set_error("plan: " + P_.error);
    return;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    set_error("no HIP device available");
    return;
  }
  ISG_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  {
    // this device's copy of the static scratch: its address, once per device
    int dev = 0;
    ISG_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (dev >= 0 && dev < kMaxDevices) {
      if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e == hipSuccess && p)
          g_scratch_ptr[dev] = p;
        else
          set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      }
      scratch_ = g_scratch_ptr[dev];
    }
  }
  int L = P_.L, R = cfg_.max_rhs;
  // multipole and local stores, all levels, [level][rhs][4^i S_i]
  lvl_off_.assign(L, 0);
  lvl_size_.assign(L, 0);
  size_t tot = 0;
  for (int i = 0; i < L; i++) {
    lvl_size_[i] = (size_t)P_.nclus[i] * P_.nsamp[i];
    lvl_off_[i] = tot;
    tot += lvl_size_[i] * R;
  }
  has32_ = cfg_.precision == PREC_MIXED;
  if (has32_) {
    // The nearest far-field translation of level i (boxes two apart, x = 2 k0 a_i)
    // runs to order P_i; once P_i exceeds x its terms grow like |Y_P(x)|, the
    // locals carry components that cancel in L2P, and FP32 rounding (6e-8) is
    // amplified by that factor (43 at the default beta 5, 2.1e4 at beta 9).
    // The FP32 operator is refused where that exceeds its 1e-5 accuracy.
    double amp = 0;
    for (int i = 2; i < L; i++) amp = std::max(amp, std::fabs(yn(P_.nterm[i], 2 * P_.k0 * P_.csize[i])));
    if (!(6e-8 * amp <= 1e-5)) {
      set_error("precision 1: the FP32 operator would lose accuracy at this beta (translation terms up to " +
                std::to_string(amp) + " amplify FP32 rounding beyond 1e-5); use precision 0 or beta <= 6");
      return;
    }
  }
  upload(op64_);
  if (has32_) {
    upload(op32_);
    build_split16();
  }
  auto up_i = [&](const std::vector<int>& src) -> int* {
    int* p = (int*)dalloc(src.size() * sizeof(int));
    if (p) ISG_HIP(hipMemcpy(p, src.data(), src.size() * sizeof(int), hipMemcpyHostToDevice));
    return p;
  };
  interp0_.assign(L, nullptr);
  anterp0_.assign(L, nullptr);
  for (int i = 0; i < L; i++) {
    if (!P_.interp0[i].empty()) interp0_[i] = up_i(P_.interp0[i]);
    if (!P_.anterp0[i].empty()) anterp0_[i] = up_i(P_.anterp0[i]);
  }
  rx0_ = up_i(P_.rx0);
  rxaI_ = up_i(P_.rxaI);
  plan_interp(op64_, m2m64_, l2l64_);
  if (has32_) plan_interp(op32_, m2m32_, l2l32_);
  for (auto& row : tune_)
    for (auto& t : row) {
      t.m2l.assign(P_.L, -1);
      t.m2m.assign(P_.L, -1);
      t.l2l.assign(P_.L, -1);
    }
  vec_.assign(V_COUNT, nullptr);
  for (int k = 0; k < V_COUNT; k++) vec_[k] = (zd*)dalloc((size_t)R * P_.N * sizeof(zd));
  vec32_.assign(V_COUNT, nullptr);
  if (has32_)
    for (int k = 0; k < V_COUNT; k++) vec32_[k] = (cplx<float>*)dalloc((size_t)R * P_.N * sizeof(cplx<float>));
  nblk_ = (int)std::min<long>(1024, std::max<long>(1, (P_.N + kBlock * 4 - 1) / (kBlock * 4)));
  // partial sums of the vector kernels (nblk_ workgroups) and of the
  // near-field epilogues (one per 16 leaf boxes at most); the host copy takes
  // either (the residual check reads the near-field ones)
  part_blocks_ = std::max<long>(nblk_, P_.B / 16);
  part_ = (double2*)dalloc((size_t)R * kSlots * part_blocks_ * sizeof(double2));
  state_ = (BiState*)dalloc((size_t)R * sizeof(BiState));
  active_ = (int*)dalloc((size_t)R * sizeof(int));
  tolv_ = (double*)dalloc((size_t)R * sizeof(double));
  lam_ = (zd*)dalloc((size_t)R * sizeof(zd));
  bratio_ = (double*)dalloc((size_t)R * sizeof(double));
  ISG_HIP(hipHostMalloc((void**)&active_h_, (size_t)R * sizeof(int), hipHostMallocDefault));
  ISG_HIP(hipHostMalloc((void**)&state_h_, (size_t)R * sizeof(BiState), hipHostMallocDefault));
  ISG_HIP(hipHostMalloc((void**)&tol_h_, (size_t)R * sizeof(double), hipHostMallocDefault));
  ISG_HIP(hipHostMalloc((void**)&part_h_, (size_t)R * kSlots * part_blocks_ * sizeof(double2), hipHostMallocDefault));
  rn_.reserve(R);
  bn_.reserve(R);
  if (ok()) {
    ISG_HIP(hipMemset(state_, 0, (size_t)R * sizeof(BiState)));
    ISG_HIP(hipMemset(active_, 0, (size_t)R * sizeof(int)));
  }
  if (ok() && cfg_.plan.radius > 0) setup_receivers();
}

// Operator set-up on the device (setup_mlfma.cpp:453-631): near blocks,
// P2M/L2P patterns, shifts and translation operators computed by kernels into
// the FP64 tables; the Hankel functions of each level's 40 translation
// offsets come from one thread per offset (isg_setup_kernels.hpp).
void Engine::device_tables(DevOp<double>& d) {
  const int L = P_.L, Sl = P_.nsamp[L - 1];
  const double k0 = P_.k0, res = cfg_.plan.res;
  d.near = (zd*)dalloc((size_t)9 * 64 * 64 * sizeof(zd));
  d.coeffT = (zd*)dalloc((size_t)64 * Sl * sizeof(zd));
  d.basisT = (zd*)dalloc((size_t)Sl * 64 * sizeof(zd));
  for (int i = 0; i < L - 1; i++) {
    d.shiftM[i] = (zd*)dalloc((size_t)4 * P_.nsamp[i] * sizeof(zd));
    d.shiftL[i] = (zd*)dalloc((size_t)4 * P_.nsamp[i] * sizeof(zd));
  }
  int Pmax = 0, Mmax = 0;
  std::vector<int> Mlev(L, 0);
  for (int i = 2; i < L; i++) {
    d.trans[i] = (zd*)dalloc((size_t)49 * P_.nsamp[i] * sizeof(zd));
    // backward-recurrence start: past the largest order and argument of the level
    const double xmax = k0 * P_.csize[i] * std::sqrt(18.0);
    const int top = std::max(P_.nterm[i], (int)std::ceil(xmax));
    Mlev[i] = top + 32 + (int)std::sqrt(40.0 * top);
    Pmax = std::max(Pmax, P_.nterm[i]);
    Mmax = std::max(Mmax, Mlev[i]);
  }
  zd* h = (zd*)dalloc((size_t)49 * (Pmax + 1) * sizeof(zd));
  double* jw = (double*)dalloc((size_t)49 * (Mmax + 2) * sizeof(double));
  if (!ok()) return;
  setup::k_near_table<<<(9 * 64 * 64 + 255) / 256, 256, 0, stream_>>>(k0, res, d.near);
  setup::k_p2m_table<<<(64 * Sl + 255) / 256, 256, 0, stream_>>>(k0, res, P_.csize[L - 1], Sl, d.coeffT, d.basisT);
  for (int i = 0; i < L - 1; i++)
    setup::k_shift_table<<<(4 * P_.nsamp[i] + 255) / 256, 256, 0, stream_>>>(k0, P_.csize[i] / 4, P_.nsamp[i],
                                                                            d.shiftM[i], d.shiftL[i]);
  for (int i = 2; i < L; i++) {
    const int S = P_.nsamp[i], Pn = P_.nterm[i];
    setup::k_hankel<<<1, 64, 0, stream_>>>(k0, P_.csize[i], Pn, Mlev[i], jw, h);
    setup::k_trans_table<<<dim3((S + 255) / 256, 49), 256, 0, stream_>>>(S, Pn, P_.csize[i], h, d.trans[i]);
  }
  ISG_HIP(hipGetLastError());
  ISG_HIP(hipStreamSynchronize(stream_));
}

// Finite-radius receivers: the numrx x N operator on the device (FP64) and a
// rocBLAS handle on the engine stream.
void Engine::setup_receivers() {
  const int Rx = cfg_.plan.numrx, R = rx_rhs();
  const long N = P_.N;
  rxm_ = (zd*)dalloc((size_t)Rx * N * sizeof(zd));
  zd* pos = (zd*)dalloc((size_t)N * sizeof(zd));
  rx_kc_ = (int)std::min<long>(N, 16384);  // N is 64 * 4^k: a multiple of the chunk
  rxpart_ = (zd*)dalloc((size_t)(N / rx_kc_) * Rx * R * sizeof(zd));
  rxc_ = (zd*)dalloc((size_t)Rx * R * sizeof(zd));
  rxin_ = (zd*)dalloc((size_t)N * R * sizeof(zd));
  if (!ok()) return;
  ISG_HIP(hipMemcpy(pos, P_.pos.data(), (size_t)N * sizeof(zd), hipMemcpyHostToDevice));
  k_rx_matrix<<<dim3(nblocks(N), Rx), dim3(kBlock), 0, stream_>>>(pos, N, Rx, rx_kc_, cfg_.plan.radius, P_.k0,
                                                                   cfg_.plan.res * cfg_.plan.res, rxm_);
  ISG_HIP(hipGetLastError());
  rocblas_handle h = nullptr;
  if (rocblas_create_handle(&h) != rocblas_status_success || rocblas_set_stream(h, stream_) != rocblas_status_success) {
    set_error("rocBLAS handle creation failed");
    return;
  }
  blas_ = h;
  ISG_HIP(hipStreamSynchronize(stream_));
}

// r[t][m] = sum_n K[m][n] x[t][n]: x copied into the operator's chunking, one
// strided-batched ZGEMM (batch b: the Rx x kc block of chunk b times the
// kc x nrhs block of x) and a fixed-order sum of the chunks.
void Engine::rx_apply(const zd* x, zd* r, int nrhs) {
  const int Rx = cfg_.plan.numrx;
  const long N = P_.N, kc = rx_kc_;
  const int nb = (int)(N / kc);
  const rocblas_double_complex one = {1.0, 0.0}, zero = {0.0, 0.0};
  k_rx_chunk<<<nblocks(N * nrhs), kBlock, 0, stream_>>>(x, N, nrhs, kc, rxin_);
  rocblas_status st = rocblas_zgemm_strided_batched(
      (rocblas_handle)blas_, rocblas_operation_transpose, rocblas_operation_none, Rx, nrhs, (rocblas_int)kc, &one,
      (const rocblas_double_complex*)rxm_, (rocblas_int)kc, (rocblas_stride)Rx * kc,
      (const rocblas_double_complex*)rxin_, (rocblas_int)kc, (rocblas_stride)nrhs * kc, &zero,
      (rocblas_double_complex*)rxpart_, Rx, (rocblas_stride)Rx * nrhs, nb);
  if (st != rocblas_status_success) {
    set_error("rocblas_zgemm_strided_batched failed");
    return;
  }
  const long len = (long)Rx * nrhs;
  k_sum_parts<<<nblocks(len), kBlock, 0, stream_>>>(rxpart_, len, nb, r);
  ISG_HIP(hipGetLastError());
}

// x[t][n] = sum_m conj(K[m][n]) r[t][m] = conj(sum_m K[m][n] conj(r[t][m])): one
// strided-batched ZGEMM over the chunks (batch b: chunk b read as the kc x numrx
// column-major block, times the conjugated receiver values, into x[.][b kc ..]),
// conjugations before and after.
void Engine::rx_apply_h(const zd* r, zd* x, int nrhs) {
  const int Rx = cfg_.plan.numrx;
  const long N = P_.N, kc = rx_kc_;
  const int nb = (int)(N / kc);
  const rocblas_double_complex one = {1.0, 0.0}, zero = {0.0, 0.0};
  k_conj_copy<<<nblocks((long)Rx * nrhs), kBlock, 0, stream_>>>(r, rxc_, (long)Rx * nrhs);
  rocblas_status st = rocblas_zgemm_strided_batched(
      (rocblas_handle)blas_, rocblas_operation_none, rocblas_operation_none, (rocblas_int)kc, nrhs, Rx, &one,
      (const rocblas_double_complex*)rxm_, (rocblas_int)kc, (rocblas_stride)Rx * kc,
      (const rocblas_double_complex*)rxc_, Rx, (rocblas_stride)0, &zero, (rocblas_double_complex*)x,
      (rocblas_int)N, (rocblas_stride)kc, nb);
  if (st != rocblas_status_success) {
    set_error("rocblas_zgemm_strided_batched failed (adjoint receiver product)");
    return;
  }
  k_conj_copy<<<nblocks(N * nrhs), kBlock, 0, stream_>>>(x, x, N * nrhs);
  ISG_HIP(hipGetLastError());
}

Engine::~Engine() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  if (blas_) (void)rocblas_destroy_handle((rocblas_handle)blas_);
  for (auto& k : kt_) {
    for (auto e : k.ev0) (void)hipEventDestroy(e);
    for (auto e : k.ev1) (void)hipEventDestroy(e);
  }
  for (void* p : allocs_) (void)hipFree(p);
  if (active_h_) (void)hipHostFree(active_h_);
  if (state_h_) (void)hipHostFree(state_h_);
  if (tol_h_) (void)hipHostFree(tol_h_);
  if (part_h_) (void)hipHostFree(part_h_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

// Split-FP16 tables of k_near_h16 and k_p2m_h16 from the FP32 operator's
// near-field, L2P and P2M tables, each with the power-of-two scale that puts
// its largest component just below 2^14.
void Engine::build_split16() {
  if (!ok()) return;
  const int S = P_.nsamp[P_.L - 1], nlc = (S + 31) / 32, ncolp = 16 * ((S + 15) / 16);
  std::vector<zf> nf((size_t)9 * 4096), bf((size_t)S * 64), cf((size_t)64 * S);
  ISG_HIP(hipMemcpy(nf.data(), op32_.near, nf.size() * sizeof(zf), hipMemcpyDeviceToHost));
  ISG_HIP(hipMemcpy(bf.data(), op32_.basisT, bf.size() * sizeof(zf), hipMemcpyDeviceToHost));
  ISG_HIP(hipMemcpy(cf.data(), op32_.coeffT, cf.size() * sizeof(zf), hipMemcpyDeviceToHost));
  if (!ok()) return;
  auto absmax = [](const std::vector<zf>& v) {
    float m = 0;
    for (const zf& c : v) m = std::max(m, std::max(std::fabs(c.x), std::fabs(c.y)));
    return m;
  };
  near16_et_ = h16_scale_exp(absmax(nf));
  bas16_eb_ = h16_scale_exp(absmax(bf));
  p2m16_ec_ = h16_scale_exp(absmax(cf));
  near16_ = dalloc((size_t)18 * 1024 * 16);
  bas16_ = dalloc((size_t)nlc * 1024 * 16);
  p2m16_ = S <= 16 * kP2mH16MaxCB ? dalloc((size_t)2 * 16 * ncolp * 16) : nullptr;
  if (!near16_ || !bas16_) return;
  // near[s][k][m] (k = source pixel), basisT[k][m] (k = local sample), coeffT[n][s] (n = pixel)
  k_split16<<<(18 * 256 + 255) / 256, 256, 0, stream_>>>(op32_.near, 2, 4096, 64, 64, 64, 64, near16_et_, 18,
                                                          (v8h*)near16_);
  k_split16<<<(nlc * 256 + 255) / 256, 256, 0, stream_>>>(op32_.basisT, nlc, 0, S, 64, 64, 64, bas16_eb_, nlc,
                                                          (v8h*)bas16_);
  if (p2m16_)
    k_split16<<<(2 * 4 * ncolp + 255) / 256, 256, 0, stream_>>>(op32_.coeffT, 2, 0, 64, S, S, ncolp, p2m16_ec_, 2,
                                                                (v8h*)p2m16_);
  ISG_HIP(hipGetLastError());
  ISG_HIP(hipStreamSynchronize(stream_));
}

// ------------------------------------------------------- interpolation plan
template <typename T>
void Engine::plan_interp(DevOp<T>& d, std::vector<InterpShape>& m2m, std::vector<InterpShape>& l2l) {
  const int L = P_.L, ni = cfg_.plan.ninter, na = 2 * ni;
  m2m.assign(L, InterpShape());
  l2l.assign(L, InterpShape());
  d.band_m2m.assign(L, nullptr);
  d.band_l2l.assign(L, nullptr);
  d.bs0_m2m.assign(L, nullptr);
  d.bs0_l2l.assign(L, nullptr);
  const size_t cz = sizeof(cplx<T>);
  auto upload_band = [&](const std::vector<double>& A, const std::vector<int>& s0, T*& pa, int*& ps) {
    std::vector<T> tmp(A.begin(), A.end());
    pa = (T*)dalloc(tmp.size() * sizeof(T));
    ps = (int*)dalloc(s0.size() * sizeof(int));
    if (pa) ISG_HIP(hipMemcpy(pa, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice));
    if (ps) ISG_HIP(hipMemcpy(ps, s0.data(), s0.size() * sizeof(int), hipMemcpyHostToDevice));
  };
  // widest column count (4, 2 or 1 blocks of 16) whose staged tile fits in LDS
  auto pick_ncb = [&](int K, int np) {
    for (int c = 4; c >= 1; c /= 2)
      if ((size_t)16 * c * (K + 1) * cz <= kLdsLimit && (4 * np) % (16 * c) == 0) return c;
    return 0;
  };
  std::vector<double> AB;
  std::vector<int> s0;
  for (int i = 0; i + 1 < L; i++) {
    const int Sp = P_.nsamp[i], Sc = P_.nsamp[i + 1], np = P_.nclus[i];
    // M2M: rows = parent samples, K = child samples
    InterpShape& m = m2m[i];
    m.ncb = pick_ncb(Sc, np);
    if (m.ncb > 0) m.kb = build_band(Sp, Sc, ni, P_.interp0[i], P_.interpT[i], 4 * kQM2M, AB, s0);
    if (m.ncb > 0 && m.kb > 0) {
      m.lds_band = (size_t)16 * m.ncb * (Sc + 1) * cz;
      upload_band(AB, s0, d.band_m2m[i], d.bs0_m2m[i]);
    } else {
      m.ncb = 0;
    }
    m.g = (Sp <= 128) ? std::min(np, pow2_floor(std::max(1, kBlock / Sp))) : 1;
    m.kch = (Sp <= 128) ? Sp : kBlock;
    int Wm = chunk_window(P_.interp0[i], Sp, Sc, ni, m.kch);
    m.lds_chunk = (size_t)4 * m.g * std::max(Wm, 0) * cz;
    m.chunk_ok = Wm > 0 && m.lds_chunk <= kLdsLimit;
    // L2L: rows = child samples, K = parent samples
    InterpShape& l = l2l[i];
    l.ncb = pick_ncb(Sp, np);
    if (l.ncb > 0) l.kb = build_band(Sc, Sp, na, P_.anterp0[i + 1], P_.anterpT[i + 1], 4 * kQL2L, AB, s0);
    if (l.ncb > 0 && l.kb > 0) {
      l.lds_band = (size_t)16 * l.ncb * (Sp + 1) * cz;
      upload_band(AB, s0, d.band_l2l[i], d.bs0_l2l[i]);
    } else {
      l.ncb = 0;
    }
    l.g = 1;
    l.kch = std::min(Sc, 64);
    int Wl = chunk_window(P_.anterp0[i + 1], Sc, Sp, na, l.kch);
    l.lds_chunk = (size_t)4 * std::max(Wl, 0) * cz;
    l.chunk_ok = Wl > 0 && l.lds_chunk <= kLdsLimit;
  }
}

// kinds: 0 band MFMA with the widest column count, 1 chunked LDS, 2 direct,
// 3 / 4 band MFMA with 16 / 32 columns per workgroup (less LDS, more waves)
bool Engine::kind_ok(const InterpShape& sh, int kind) const {
  return (kind == 0 && sh.ncb > 0) || (kind == 1 && sh.chunk_ok) || kind == 2 || (kind == 3 && sh.ncb >= 1) ||
         (kind == 4 && sh.ncb >= 2);
}

// band GEMM when it yields at least one workgroup per CU, else chunked, else direct
int Engine::interp_kind(const InterpShape& sh, int np, int nrhs) const {
  if (sh.ncb > 0 && (long)(4 * np / (16 * sh.ncb)) * nrhs >= 256) return 0;
  if (sh.chunk_ok) return 1;
  return sh.ncb > 0 ? 0 : 2;
}

// --------------------------------------------------------------- launches
// leaf radiation patterns on the matrix cores
template <typename T, typename TI>
void Engine::launch_p2m(const cplx<TI>* y, int nrhs, const int* act, int force, const NearPost<TI>* prep) {
  const int Lf = P_.L - 1, S = P_.nsamp[Lf], B = P_.B;
  const int tuned = tune(sizeof(T) == 4 ? PREC_MIXED : PREC_FP64, nrhs).p2m_nb;
  int nb = force > 0 ? force : (tuned > 0 ? tuned : (B >= 32 ? 32 : 16));
  DevOp<T>& d = ops<T>();
  cplx<T>* out = d.agm + lvl_off_[Lf];
  long as = (long)lvl_size_[Lf];
  if (nb == kNearSplit) {
    // split-FP16 P2M: FP32 operator on FP32 vectors, 32-box tiles of two
    // right-hand sides (below 4 the workgroups are few and half idle)
    if constexpr (sizeof(T) == 4 && sizeof(TI) == 4) {
      if (!prep && B % 32 == 0 && p2m16_ && (nrhs >= 4 || force == kNearSplit)) {
        dim3 grid(B / 32, (nrhs + 1) / 2), blk(kBlock);
        k_p2m_h16<<<grid, blk, 0, stream_>>>(y, P_.N, out, as, (const v8h*)p2m16_, p2m16_ec_, S, nrhs, act);
        return;
      }
    }
    nb = B >= 32 ? 32 : 16;
  }
  if (nb == 64 && sizeof(T) != 4) nb = 32;  // 64-box tiles: FP32 only (LDS)
  if (B % nb != 0) nb = 16;
  dim3 grid(B / nb, nrhs), blk(kBlock);
  if (prep) {  // residual check: y is the iterate, contrast-weighted while loading (FP64 only)
    if constexpr (sizeof(T) == 8 && sizeof(TI) == 8) {
      if (nb == 32)
        k_p2m_mfma<T, TI, 32, true><<<grid, blk, 0, stream_>>>(y, P_.N, out, as, d.coeffT, S, act, prep->g, prep->gs,
                                                               prep->herm);
      else
        k_p2m_mfma<T, TI, 16, true><<<grid, blk, 0, stream_>>>(y, P_.N, out, as, d.coeffT, S, act, prep->g, prep->gs,
                                                               prep->herm);
    } else {
      set_error("P2M: contrast weighting on load is FP64 only");
    }
    return;
  }
  if (nb == 64) {
    if constexpr (sizeof(T) == 4) k_p2m_mfma<T, TI, 64><<<grid, blk, 0, stream_>>>(y, P_.N, out, as, d.coeffT, S, act);
  } else if (nb == 32)
    k_p2m_mfma<T, TI, 32><<<grid, blk, 0, stream_>>>(y, P_.N, out, as, d.coeffT, S, act);
  else
    k_p2m_mfma<T, TI, 16><<<grid, blk, 0, stream_>>>(y, P_.N, out, as, d.coeffT, S, act);
}

// child level i+1 -> parent level i
template <typename T>
void Engine::launch_m2m(int i, int nrhs, const int* act, int force) {
  const int Sp = P_.nsamp[i], Sc = P_.nsamp[i + 1], np = P_.nclus[i], nt = cfg_.plan.ninter;
  DevOp<T>& d = ops<T>();
  cplx<T>*child = d.agm + lvl_off_[i + 1], *par = d.agm + lvl_off_[i];
  long cs = (long)lvl_size_[i + 1], ps = (long)lvl_size_[i];
  const InterpShape& sh = (sizeof(T) == 4 ? m2m32_ : m2m64_)[i];
  dim3 blk(kBlock);
  const int tuned = tune(sizeof(T) == 4 ? PREC_MIXED : PREC_FP64, nrhs).m2m[i];
  const int kind = kind_ok(sh, force) ? force : (kind_ok(sh, tuned) ? tuned : interp_kind(sh, np, nrhs));
  if (kind == 0 || kind >= 3) {
    const int ncb = kind == 3 ? 1 : (kind == 4 ? 2 : sh.ncb);
    const size_t lds = (size_t)16 * ncb * (Sc + 1) * sizeof(cplx<T>);
    dim3 grid(4 * np / (16 * ncb), nrhs);
    const T* ab = d.band_m2m[i];
    const int* s0 = d.bs0_m2m[i];
    if (ncb == 4)
      k_interp_band<T, 0, 4, kQM2M><<<grid, blk, lds, stream_>>>(child, cs, par, ps, ab, s0, sh.kb, d.shiftM[i], Sp,
                                                                 Sc, 0, act);
    else if (ncb == 2)
      k_interp_band<T, 0, 2, kQM2M><<<grid, blk, lds, stream_>>>(child, cs, par, ps, ab, s0, sh.kb, d.shiftM[i], Sp,
                                                                 Sc, 0, act);
    else
      k_interp_band<T, 0, 1, kQM2M><<<grid, blk, lds, stream_>>>(child, cs, par, ps, ab, s0, sh.kb, d.shiftM[i], Sp,
                                                                 Sc, 0, act);
  } else if (kind == 1) {
    int nch = (Sp + sh.kch - 1) / sh.kch;
    dim3 grid(np / sh.g * nch, nrhs);
    k_m2m_chunk<T><<<grid, blk, sh.lds_chunk, stream_>>>(child, cs, par, ps, d.interpT[i], interp0_[i], d.shiftM[i],
                                                         Sp, Sc, sh.g, sh.kch, nch, nt, act);
  } else {
    k_m2m<T><<<dim3(nblocks((long)np * Sp), nrhs), blk, 0, stream_>>>(child, cs, par, ps, d.interpT[i], interp0_[i],
                                                                       d.shiftM[i], Sp, Sc, np, nt, act);
  }
}

// parent level i -> child level i+1, accumulating (matvec) or assigning
// (receiver back-projection).
template <typename T>
void Engine::launch_l2l(int i, bool accumulate, int nrhs, const int* act, int force) {
  const int Sp = P_.nsamp[i], Sc = P_.nsamp[i + 1], np = P_.nclus[i], nt = 2 * cfg_.plan.ninter;
  DevOp<T>& d = ops<T>();
  cplx<T>*par = d.agl + lvl_off_[i], *child = d.agl + lvl_off_[i + 1];
  long ps = (long)lvl_size_[i], cs = (long)lvl_size_[i + 1];
  const InterpShape& sh = (sizeof(T) == 4 ? l2l32_ : l2l64_)[i];
  dim3 blk(kBlock);
  int acc = accumulate ? 1 : 0;
  const int tuned = tune(sizeof(T) == 4 ? PREC_MIXED : PREC_FP64, nrhs).l2l[i];
  const int kind = kind_ok(sh, force) ? force : (kind_ok(sh, tuned) ? tuned : interp_kind(sh, np, nrhs));
  if (kind == 0 || kind >= 3) {
    const int ncb = kind == 3 ? 1 : (kind == 4 ? 2 : sh.ncb);
    const size_t lds = (size_t)16 * ncb * (Sp + 1) * sizeof(cplx<T>);
    dim3 grid(4 * np / (16 * ncb), nrhs);
    const T* ab = d.band_l2l[i];
    const int* s0 = d.bs0_l2l[i];
    if (ncb == 4)
      k_interp_band<T, 1, 4, kQL2L><<<grid, blk, lds, stream_>>>(par, ps, child, cs, ab, s0, sh.kb, d.shiftL[i], Sc,
                                                                 Sp, acc, act);
    else if (ncb == 2)
      k_interp_band<T, 1, 2, kQL2L><<<grid, blk, lds, stream_>>>(par, ps, child, cs, ab, s0, sh.kb, d.shiftL[i], Sc,
                                                                 Sp, acc, act);
    else
      k_interp_band<T, 1, 1, kQL2L><<<grid, blk, lds, stream_>>>(par, ps, child, cs, ab, s0, sh.kb, d.shiftL[i], Sc,
                                                                 Sp, acc, act);
  } else if (kind == 1) {
    int nch = (Sc + sh.kch - 1) / sh.kch;
    dim3 grid(np * nch, nrhs);
    k_l2l_chunk<T><<<grid, blk, sh.lds_chunk, stream_>>>(par, ps, child, cs, d.anterpT[i + 1], anterp0_[i + 1],
                                                         d.shiftL[i], Sp, Sc, sh.kch, nch, nt, acc, act);
  } else {
    k_l2l<T><<<dim3(nblocks((long)4 * np * Sc), nrhs), blk, 0, stream_>>>(par, ps, child, cs, d.anterpT[i + 1],
                                                                           anterp0_[i + 1], d.shiftL[i], Sp, Sc, np,
                                                                           nt, acc, act);
  }
}

// M2L variants: TS x TS target boxes and KC samples per workgroup; 0-5
// k_m2l_tile, 6-8 k_m2l_reg (operators in registers), 9 k_m2l_par (one
// thread per parent and sample, sources in registers)
constexpr int kM2LVariants = 10;
constexpr int kM2LShape[kM2LVariants][2] = {{4, 8}, {8, 8}, {8, 16}, {16, 8}, {8, 32}, {16, 16},
                                            {8, 16}, {8, 32}, {8, 8}, {8, 16}};

template <typename T>
bool Engine::m2l_ok(int v, int i) const {
  if (v < 0 || v >= kM2LVariants) return false;
  const int TS = kM2LShape[v][0], KC = kM2LShape[v][1];
  const size_t lds = (size_t)((TS + 4) * (TS + 4) + 49) * KC * sizeof(cplx<T>);
  return (1 << i) >= TS && lds <= kLdsLimit;
}

// translations of level i.  Built-in rule: 8x8 tiles of 8 samples, 4x4 when
// 8x8 tiles would leave fewer than 1024 workgroups.
template <typename T>
void Engine::launch_m2l(int i, int nrhs, const int* act, int force) {
  const int S = P_.nsamp[i], nb = 1 << i;
  int v = force;
  if (!m2l_ok<T>(v, i)) v = tune(sizeof(T) == 4 ? PREC_MIXED : PREC_FP64, nrhs).m2l[i];
  if (!m2l_ok<T>(v, i)) v = (nb >= 8 && (long)(nb * nb / 64) * ((S + 7) / 8) * nrhs >= 1024) ? 1 : 0;
  const int TS = kM2LShape[v][0], KC = kM2LShape[v][1];
  const int nchunk = (S + KC - 1) / KC;
  const int ntile = nb * nb / (TS * TS);
  dim3 grid(ntile * nchunk, nrhs), blk(kBlock);
  DevOp<T>& d = ops<T>();
  const cplx<T>* m = d.agm + lvl_off_[i];
  cplx<T>* l = d.agl + lvl_off_[i];
  long st = (long)lvl_size_[i];
  switch (v) {
    case 0: k_m2l_tile<T, 4, 8><<<grid, blk, 0, stream_>>>(m, l, st, d.trans[i], i, S, act); break;
    case 1: k_m2l_tile<T, 8, 8><<<grid, blk, 0, stream_>>>(m, l, st, d.trans[i], i, S, act); break;
    case 2: k_m2l_tile<T, 8, 16><<<grid, blk, 0, stream_>>>(m, l, st, d.trans[i], i, S, act); break;
    case 3: k_m2l_tile<T, 16, 8><<<grid, blk, 0, stream_>>>(m, l, st, d.trans[i], i, S, act); break;
    case 6: k_m2l_reg<T, 16><<<grid, blk, 0, stream_>>>(m, l, st, d.trans[i], i, S, act); break;
    case 8: k_m2l_reg<T, 8><<<grid, blk, 0, stream_>>>(m, l, st, d.trans[i], i, S, act); break;
    case 9: k_m2l_par<T><<<grid, blk, 0, stream_>>>(m, l, st, d.trans[i], i, S, act); break;
    case 7:
      if constexpr (sizeof(T) == 4) k_m2l_reg<T, 32><<<grid, blk, 0, stream_>>>(m, l, st, d.trans[i], i, S, act);
      break;
    default:
      if constexpr (sizeof(T) == 4) {
        if (v == 4)
          k_m2l_tile<T, 8, 32><<<grid, blk, 0, stream_>>>(m, l, st, d.trans[i], i, S, act);
        else
          k_m2l_tile<T, 16, 16><<<grid, blk, 0, stream_>>>(m, l, st, d.trans[i], i, S, act);
      }
      break;
  }
}

// k_m2l_leaf: leaf M2L plus the band L2L from level Lf-1, whose 16-row blocks
// coincide with the kernel's 16-sample chunks.  Needs a leaf level of at
// least 8x8 boxes with a parent level below it that has M2L (>= 2), the band
// form of that L2L, and the LDS of both stages within the limit.
template <typename T>
bool Engine::leaf_fuse_ok() const {
  const int Lf = P_.L - 1;
  if (Lf < 3) return false;
  const InterpShape& sh = (sizeof(T) == 4 ? l2l32_ : l2l64_)[Lf - 1];
  // kb <= 96: the kernel's register window
  if (sh.ncb <= 0 || sh.kb <= 0 || sh.kb > P_.nsamp[Lf - 1] || sh.kb > 96) return false;
  const size_t dyn = std::max<size_t>(16 * (size_t)sh.kb, 64 * kParKC) + 4 * (size_t)sh.kb;
  const size_t lds = ((size_t)(kParSB * kParSB + 49) * kParKC + dyn) * sizeof(cplx<T>);
  return lds <= kLdsLimit;
}

template <typename T>
void Engine::launch_m2l_leaf(int nrhs, const int* act) {
  const int Lf = P_.L - 1, S = P_.nsamp[Lf], Sp = P_.nsamp[Lf - 1], nb = 1 << Lf;
  const InterpShape& sh = (sizeof(T) == 4 ? l2l32_ : l2l64_)[Lf - 1];
  DevOp<T>& d = ops<T>();
  const int nchunk = (S + kParKC - 1) / kParKC;
  dim3 grid(nb * nb / (kParTS * kParTS) * nchunk, nrhs), blk(kBlock);
  const size_t lds = (std::max<size_t>(16 * (size_t)sh.kb, 64 * kParKC) + 4 * (size_t)sh.kb) * sizeof(cplx<T>);
  k_m2l_leaf<T><<<grid, blk, lds, stream_>>>(d.agm + lvl_off_[Lf], d.agl + lvl_off_[Lf], (long)lvl_size_[Lf],
                                             d.trans[Lf], Lf, S, d.agl + lvl_off_[Lf - 1], (long)lvl_size_[Lf - 1],
                                             Sp, d.band_l2l[Lf - 1], d.bs0_l2l[Lf - 1], sh.kb, d.shiftL[Lf - 1], act);
}

// Near field + L2P on the matrix cores (k_near_mfma), or L2P alone when
// with_near is false (receiver back-projection).
template <typename T, typename TI>
int Engine::launch_near(const cplx<TI>* y, cplx<TI>* r, int nrhs, const int* act, bool with_near, int force,
                        const NearPost<TI>* post) {
  const int Lf = P_.L - 1, S = P_.nsamp[Lf], B = P_.B;
  const long N = P_.N, ls = (long)lvl_size_[Lf];
  const int tuned = tune(sizeof(T) == 4 ? PREC_MIXED : PREC_FP64, nrhs).near_nb;
  int nb = force > 0 ? force : (tuned > 0 ? tuned : (B >= 32 ? 32 : 16));
  DevOp<T>& d = ops<T>();
  const cplx<T>* loc = d.agl + lvl_off_[Lf];
  const NearPost<TI> pp = post ? *post : NearPost<TI>();
  const int kt = sizeof(T) == 4 ? KT_NEAR32 : KT_NEAR;
  if (nb == kNearSplitWide) {
    // split-FP16 kernel with 128-box tiles of one right-hand side
    if constexpr (sizeof(T) == 4 && sizeof(TI) == 4) {
      if (with_near && B % 128 == 0 && near16_ && bas16_ && pp.mode != 3) {
        dim3 grid(B / 128, nrhs), blk(kBlock);
        auto go = [&](auto modec) {
          k_near_h16w<decltype(modec)::value><<<grid, blk, kH16WLds, stream_>>>(
              y, N, loc, ls, r, N, (const v8h*)near16_, near16_et_, (const v8h*)bas16_, bas16_eb_, S, 1 << Lf, act,
              pp);
        };
        prof_begin(kt);
        if (pp.mode == 1)
          go(std::integral_constant<int, 1>());
        else if (pp.mode == 2)
          go(std::integral_constant<int, 2>());
        else
          go(std::integral_constant<int, 0>());
        prof_end(kt);
        return B / 128;
      }
    }
    nb = kNearSplit;
  }
  if (nb == kNearSplit) {
    // split-FP16 kernel: FP32 operator on FP32 vectors, near field included,
    // 64-box tiles, BiCGSTAB post steps (not the FP64 residual check)
    if constexpr (sizeof(T) == 4 && sizeof(TI) == 4) {
      // (pairs of right-hand sides: below 4 the workgroups are few and half idle)
      if (with_near && B % 64 == 0 && near16_ && bas16_ && pp.mode != 3 && (nrhs >= 4 || force == kNearSplit)) {
        dim3 grid(B / 64, (nrhs + 1) / 2), blk(kBlock);
        auto go = [&](auto modec) {
          k_near_h16<decltype(modec)::value><<<grid, blk, kH16Lds, stream_>>>(
              y, N, loc, ls, r, N, (const v8h*)near16_, near16_et_, (const v8h*)bas16_, bas16_eb_, S, 1 << Lf, nrhs,
              act, pp);
        };
        prof_begin(kt);
        if (pp.mode == 1)
          go(std::integral_constant<int, 1>());
        else if (pp.mode == 2)
          go(std::integral_constant<int, 2>());
        else
          go(std::integral_constant<int, 0>());
        prof_end(kt);
        return B / 64;
      }
    }
    nb = B >= 32 ? 32 : 16;
  }
  if (nb == 64 && sizeof(T) != 4) nb = 32;  // 64-box tiles: FP32 only (LDS)
  if (B % nb != 0) nb = 16;
  size_t lds = (size_t)near_lds_elems(nb, S) * sizeof(cplx<T>);
  dim3 grid(B / nb, nrhs), blk(kBlock);
  if (pp.mode == 3 && !(sizeof(T) == 8 && sizeof(TI) == 8 && nb <= 32)) {
    set_error("near field: the residual-check epilogue is FP64 only");
    return B / nb;
  }
  const int wn = with_near ? 1 : 0, lv = 1 << Lf;
  // one instantiation per post mode: the plain product keeps the register
  // budget (and occupancy) of a kernel without the post epilogue
  auto go = [&](auto nbc, auto modec) {
    k_near_mfma<T, TI, decltype(nbc)::value, decltype(modec)::value>
        <<<grid, blk, lds, stream_>>>(y, N, loc, ls, r, N, d.near, d.basisT, S, lv, wn, act, pp);
  };
  using I16 = std::integral_constant<int, 16>;
  using I32 = std::integral_constant<int, 32>;
  using I64 = std::integral_constant<int, 64>;
  using M0 = std::integral_constant<int, 0>;
  using M1 = std::integral_constant<int, 1>;
  using M2 = std::integral_constant<int, 2>;
  using M3 = std::integral_constant<int, 3>;
  if (with_near) prof_begin(kt);
  if (nb == 64) {
    if constexpr (sizeof(T) == 4) {
      if (pp.mode == 1)
        go(I64(), M1());
      else if (pp.mode == 2)
        go(I64(), M2());
      else
        go(I64(), M0());
    }
  } else if (nb == 32) {
    if (pp.mode == 1)
      go(I32(), M1());
    else if (pp.mode == 2)
      go(I32(), M2());
    else if (pp.mode == 3) {
      if constexpr (sizeof(T) == 8 && sizeof(TI) == 8) go(I32(), M3());
    } else
      go(I32(), M0());
  } else {
    if (pp.mode == 1)
      go(I16(), M1());
    else if (pp.mode == 2)
      go(I16(), M2());
    else if (pp.mode == 3) {
      if constexpr (sizeof(T) == 8 && sizeof(TI) == 8) go(I16(), M3());
    } else
      go(I16(), M0());
  }
  if (with_near) prof_end(kt);
  return B / nb;
}

// r = G y (y already contrast-weighted).  Stage order of matvec_gpu.cu:103-264:
// P2M, M2M up to level 2, M2L on levels 2..leaf, L2L down, then near field
// and L2P together.
template <typename T, typename TI>
void Engine::apply(const cplx<TI>* y, cplx<TI>* r, int nrhs, const int* act, const NearPost<TI>* post) {
  const int L = P_.L, Lf = L - 1;
  // post mode 3 (residual check): y is the iterate, weighted while P2M and the near field load it
  prof_begin(KT_P2M);
  launch_p2m<T, TI>(y, nrhs, act, -1, (post && post->mode == 3) ? post : nullptr);
  prof_end(KT_P2M);
  prof_begin(KT_INTERP);
  for (int i = L - 2; i >= 2; i--) launch_m2m<T>(i, nrhs, act);
  prof_end(KT_INTERP);
  const int kt = sizeof(T) == 4 ? KT_M2L32 : KT_M2L;
  // with the fused leaf kernel the last L2L (level Lf-1 -> Lf) runs inside the
  // leaf M2L, after the L2L passes above it
  const bool fuse = tune(sizeof(T) == 4 ? PREC_MIXED : PREC_FP64, nrhs).leaf == 1 && leaf_fuse_ok<T>();
  const int mtop = fuse ? Lf - 1 : Lf, ltop = fuse ? L - 3 : L - 2;
  prof_begin(kt);
  for (int i = 2; i <= mtop; i++) launch_m2l<T>(i, nrhs, act);
  prof_end(kt);
  prof_begin(KT_INTERP, false);  // timed with the M2M launch set of this product
  for (int i = 2; i <= ltop; i++) launch_l2l<T>(i, true, nrhs, act);
  prof_end(KT_INTERP);
  if (fuse) {
    prof_begin(kt, false);  // timed with the M2L launch set of this product
    launch_m2l_leaf<T>(nrhs, act);
    prof_end(kt);
  }
  near_blocks_ = launch_near<T, TI>(y, r, nrhs, act, true, -1, post);
}

void Engine::apply_mlfma(const zd* y, zd* r, int nrhs, const int* act, int prec) { apply_vec<double>(y, r, nrhs, act, prec); }

template <typename V>
cplx<V>* Engine::vb(int k) {
  if constexpr (sizeof(V) == 4)
    return vec32_[k];
  else
    return vec_[k];
}

// FP32 vectors always go with the FP32 operator; FP64 vectors with either
template <typename V>
void Engine::apply_vec(const cplx<V>* y, cplx<V>* r, int nrhs, const int* act, int prec, const NearPost<V>* post) {
  if (!ok()) return;
  if constexpr (sizeof(V) == 4) {
    apply<float, float>(y, r, nrhs, act, post);
  } else {
    if (prec == PREC_MIXED && has32_)
      apply<float, double>(y, r, nrhs, act, post);
    else
      apply<double, double>(y, r, nrhs, act, post);
  }
  ISG_HIP(hipGetLastError());
}

void Engine::mlfma_f32(const cplx<float>* x, cplx<float>* r, int nrhs) {
  if (!has32_) {
    set_error("mlfma_f32: the context has no FP32 operator (precision 1 needed)");
    return;
  }
  apply_vec<float>(x, r, nrhs, nullptr, PREC_MIXED);
  matvecs += nrhs;
  matvecs32 += nrhs;
}

void Engine::mlfma(const zd* x, zd* r, int nrhs, const int* act, int prec) {
  apply_mlfma(x, r, nrhs, act, prec);
  matvecs += nrhs;
  if (prec == PREC_MIXED && has32_) matvecs32 += nrhs;
}

// matvec_gpu.cu:266-368 with the top levels and the receiver interpolation
// on the device as well (FP64).
void Engine::aggregate(const zd* x, zd* r, int nrhs) {
  if (!ok()) return;
  prof_begin(KT_RX);
  if (cfg_.plan.radius > 0) {
    rx_apply(x, r, nrhs);
  } else {
    const int L = P_.L;
    launch_p2m<double, double>(x, nrhs, nullptr);
    for (int i = L - 2; i >= 0; i--) launch_m2m<double>(i, nrhs, nullptr);
    int R = cfg_.plan.numrx;
    k_rx_interp<double><<<dim3(nblocks(R), nrhs), dim3(kBlock), 0, stream_>>>(
        op64_.agm + lvl_off_[0], (long)lvl_size_[0], r, R, op64_.rxT, rx0_, P_.nsamp[0], R, cfg_.plan.ninter);
    ISG_HIP(hipGetLastError());
  }
  prof_end(KT_RX);
}

// matvec_gpu.cu:369-470 (argument order out x, in r as in the reference),
// every level on the device (FP64).
void Engine::aggregateh(const zd* r, zd* x, int nrhs) {
  if (!ok()) return;
  prof_begin(KT_RX);
  if (cfg_.plan.radius > 0) {
    rx_apply_h(r, x, nrhs);
  } else {
    const int L = P_.L;
    int R = cfg_.plan.numrx;
    double sc = (double)R * cfg_.plan.res * cfg_.plan.res;
    k_rx_anterp<double><<<dim3(nblocks(P_.nsamp[0]), nrhs), dim3(kBlock), 0, stream_>>>(
        r, R, op64_.agl + lvl_off_[0], (long)lvl_size_[0], op64_.rxaT, rxaI_, P_.nsamp[0], 2 * cfg_.plan.ninter,
        sc);
    for (int i = 0; i <= L - 2; i++) launch_l2l<double>(i, false, nrhs, nullptr);
    launch_near<double, double>(nullptr, x, nrhs, nullptr, false);
    ISG_HIP(hipGetLastError());
  }
  prof_end(KT_RX);
}

void Engine::launch_prep(const zd* x, const zd* g, long gs, zd* y, bool herm, int nrhs) {
  k_prep<<<dim3(nblk_, nrhs), dim3(kBlock), 0, stream_>>>(x, g, gs, y, P_.N, herm ? 1 : 0);
}

// ---- live kernel timing (HIP events on the engine stream)
void Engine::profile(bool on) {
  if (stream_) ISG_HIP(hipStreamSynchronize(stream_));
  harvest();
  prof_ = on;
  for (auto& k : kt_) {
    k.ms = 0;
    k.launches = 0;
  }
}

void Engine::prof_begin(int id, bool new_set) {
  if (!prof_) return;
  KTimer& k = kt_[id];
  if (k.used == (int)k.ev0.size()) {
    if (k.used >= 512) {
      ISG_HIP(hipStreamSynchronize(stream_));
      harvest();
    } else {
      hipEvent_t a, b;
      ISG_HIP(hipEventCreate(&a));
      ISG_HIP(hipEventCreate(&b));
      k.ev0.push_back(a);
      k.ev1.push_back(b);
      k.starts.push_back(0);
    }
  }
  k.starts[k.used] = new_set ? 1 : 0;
  ISG_HIP(hipEventRecord(k.ev0[k.used], stream_));
}

void Engine::prof_end(int id) {
  if (!prof_) return;
  KTimer& k = kt_[id];
  ISG_HIP(hipEventRecord(k.ev1[k.used], stream_));
  k.used++;
}

// Accumulate the elapsed time of every recorded launch; the stream must have
// reached the recorded events (callers synchronize first).
void Engine::harvest() {
  for (auto& k : kt_) {
    for (int i = 0; i < k.used; i++) {
      float ms = 0;
      if (hipEventSynchronize(k.ev1[i]) == hipSuccess && hipEventElapsedTime(&ms, k.ev0[i], k.ev1[i]) == hipSuccess) {
        k.ms += ms;
        k.launches += k.starts[i];
      }
    }
    k.used = 0;
  }
}

void Engine::profile_read(int id, double& ms, long long& launches) {
  ISG_HIP(hipStreamSynchronize(stream_));
  harvest();
  ms = kt_[id].ms;
  launches = kt_[id].launches;
}

// ---------------------------------------------------------------- solver
// bicgs_gpu.cu:116-224, batched over right-hand sides with device-resident
// scalars and per-system tolerances (tol_h_ -> tolv_); the host reads only
// the per-system activity flags once per iteration.
template <typename V>
void Engine::run_bicgs(zd* x, const zd* g, long gs, const zd* b, bool herm, int nrhs, bool x_zero, int maxit,
                       int prec, const cplx<V>* r0) {
  const long N = P_.N;
  cplx<V>*r = vb<V>(V_R), *rt = vb<V>(V_RT), *y = vb<V>(V_Y), *op = vb<V>(V_OP);
  checked_x_ = nullptr;
  const int h = herm ? 1 : 0;
  dim3 grid(nblk_, nrhs), blk(kBlock);
  ISG_HIP(hipMemcpyAsync(tolv_, tol_h_, nrhs * sizeof(double), hipMemcpyHostToDevice, stream_));
  const long long f32 = (prec == PREC_MIXED && has32_) ? 1 : 0;
  if (!x_zero && !r0) {
    k_prep<V><<<grid, blk, 0, stream_>>>(x, g, gs, y, N, h);
    apply_vec<V>(y, op, nrhs, nullptr, prec);
    matvecs += nrhs;
    matvecs32 += f32 * nrhs;
  }
  k_bi_init<V><<<grid, blk, 0, stream_>>>(x, op, g, gs, b, r0, r, rt, N, h, x_zero ? 1 : 0, part_);
  k_bi_fin<<<dim3(nrhs), blk, 0, stream_>>>(PH_INIT, state_, active_, part_, nblk_, tolv_, maxit);
  maxit_ = maxit;
  // one iteration = bicgs_body; with graphs it is captured once per call and
  // replayed, the host reading the activity flags after each replay
  hipGraphExec_t exec = nullptr;
  if (cfg_.graph && !prof_) {
    hipGraph_t gr = nullptr;
    ISG_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    bicgs_body<V>(x, g, gs, herm, nrhs, prec);
    ISG_HIP(hipStreamEndCapture(stream_, &gr));
    if (gr) {
      ISG_HIP(hipGraphInstantiate(&exec, gr, nullptr, nullptr, 0));
      (void)hipGraphDestroy(gr);
    }
    if (!ok()) return;
  }
  ISG_HIP(hipMemcpyAsync(active_h_, active_, nrhs * sizeof(int), hipMemcpyDeviceToHost, stream_));
  for (int it = 0; it <= maxit; it++) {
    ISG_HIP(hipStreamSynchronize(stream_));
    if (!ok()) break;
    if (prof_) harvest();
    int nact = 0;
    for (int i = 0; i < nrhs; i++) nact += active_h_[i] != 0;
    if (nact == 0) break;
    if (exec)
      ISG_HIP(hipGraphLaunch(exec, stream_));
    else
      bicgs_body<V>(x, g, gs, herm, nrhs, prec);
    matvecs += 2 * nact;
    matvecs32 += f32 * 2 * nact;
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  ISG_HIP(hipMemcpyAsync(state_h_, state_, nrhs * sizeof(BiState), hipMemcpyDeviceToHost, stream_));
  ISG_HIP(hipStreamSynchronize(stream_));
}

// One BiCGSTAB iteration for every active system (bicgs_gpu.cu:146-213),
// ending with the copy of the activity flags to the host.  No memset: the
// sequence is captured into a graph.
template <typename V>
void Engine::bicgs_body(zd* x, const zd* g, long gs, bool herm, int nrhs, int prec) {
  const long N = P_.N;
  cplx<V>*r = vb<V>(V_R), *rt = vb<V>(V_RT), *p = vb<V>(V_P), *v = vb<V>(V_V), *s = vb<V>(V_S), *t = vb<V>(V_T),
     *y = vb<V>(V_Y);
  const int h = herm ? 1 : 0;
  const int maxit = maxit_;
  dim3 grid(nblk_, nrhs), blk(kBlock);
  // v = post(p, G y) and t = post(s, G y) with their dot products are the
  // near-field kernel's epilogue (NearPost); its partials are per near workgroup
  NearPost<V> post;
  post.g = g;
  post.gs = gs;
  post.herm = h;
  post.part = part_;
  k_bi_pupd<V><<<grid, blk, 0, stream_>>>(r, p, v, g, gs, y, N, h, state_);
  post.mode = 1;
  post.x = p;
  post.rt = rt;
  apply_vec<V>(y, v, nrhs, active_, prec, &post);
  k_bi_fin<<<dim3(nrhs), blk, 0, stream_>>>(PH_ALPHA, state_, active_, part_, near_blocks_, tolv_, maxit);
  k_bi_supd<V><<<grid, blk, 0, stream_>>>(r, v, g, gs, s, y, N, h, state_);
  post.mode = 2;
  post.x = s;
  post.rt = nullptr;
  apply_vec<V>(y, t, nrhs, active_, prec, &post);
  k_bi_fin<<<dim3(nrhs), blk, 0, stream_>>>(PH_OMEGA, state_, active_, part_, near_blocks_, tolv_, maxit);
  prof_begin(KT_VEC);
  k_bi_xrupd<V><<<grid, blk, 0, stream_>>>(x, p, s, t, r, rt, N, state_, part_);
  prof_end(KT_VEC);
  k_bi_fin<<<dim3(nrhs), blk, 0, stream_>>>(PH_RESID, state_, active_, part_, nblk_, tolv_, maxit);
  ISG_HIP(hipMemcpyAsync(active_h_, active_, nrhs * sizeof(int), hipMemcpyDeviceToHost, stream_));
}

// rn, bn: per-system norms of b - A x (FP64 operator) and of b; G prep(x)
// stays in vec_[V_OP] until the next product.
void Engine::true_residual(const zd* x, const zd* g, long gs, const zd* b, bool herm, int nrhs,
                           std::vector<double>& rn, std::vector<double>& bn) {
  const int h = herm ? 1 : 0;
  zd* op = vec_[V_OP];
  checked_x_ = nullptr;
  // op = G prep(x) with the FP64 operator, prep(x) formed while P2M and the
  // near field load x; the near-field kernel's epilogue also forms the partial
  // sums of |b|^2 and |b - A x|^2 (the residual vector itself is formed later,
  // from op, if a correction needs it)
  NearPost<double> pp;
  pp.mode = 3;
  pp.x = x;
  pp.b = b;
  pp.g = g;
  pp.gs = gs;
  pp.herm = h;
  pp.part = part_;
  apply_vec<double>(x, op, nrhs, nullptr, PREC_FP64, &pp);
  matvecs += nrhs;
  const int nb = near_blocks_;
  if (nb > part_blocks_) {  // cannot happen with the buffers sized above; never copy past part_h_
    set_error("residual check: more near-field workgroups than partial-sum slots");
    return;
  }
  ISG_HIP(hipMemcpyAsync(part_h_, part_, (size_t)nrhs * kSlots * nb * sizeof(double2), hipMemcpyDeviceToHost,
                         stream_));
  ISG_HIP(hipStreamSynchronize(stream_));
  rn.assign(nrhs, 0.0);
  bn.assign(nrhs, 0.0);
  for (int k = 0; k < nrhs; k++) {
    const double2* pr = part_h_ + (size_t)k * kSlots * nb;
    double sb = 0, sr = 0;
    for (int j = 0; j < nb; j++) {
      sb += pr[j].x;
      sr += pr[nb + j].x;
    }
    bn[k] = std::sqrt(sb);
    rn[k] = std::sqrt(sr);
  }
}

const zd* Engine::checked_product(const zd* x, int nrhs, bool herm) const {
  return (checked_x_ && checked_x_ == x && checked_n_ == nrhs && checked_h_ == herm) ? vec_[V_OP] : nullptr;
}

void Engine::bicgs(zd* x, const zd* g, long gs, const zd* b, bool herm, int nrhs, bool x_zero, double tol,
                   int maxit, std::vector<SolveStat>& out, int prec) {
  solve(x, g, gs, b, herm, nrhs, x_zero, tol, maxit, out, prec, false);
}

size_t Engine::static_scratch_bytes(int nrhs) const {
  size_t off = 0;
  auto take = [&](size_t bytes) { off += (bytes + 255) & ~(size_t)255; };
  take((size_t)nrhs * kSlots * part_blocks_ * sizeof(double2));
  take((size_t)nrhs * sizeof(BiState));
  take((size_t)nrhs * sizeof(int));
  take((size_t)nrhs * sizeof(double));
  take((size_t)nrhs * sizeof(zd));
  return off;
}

// The solve of bicgs() with its scratch (partials, scalars, flags,
// tolerances, blend coefficients) in this device's static array instead of
// the engine's allocations, launched kernel by kernel (no graph is
// instantiated) and without the live timers: no device memory is allocated.
// The caller has checked static_scratch_bytes(nrhs) <= kStaticScratchBytes.
void Engine::bicgs_static(zd* x, const zd* g, long gs, const zd* b, bool herm, int nrhs, bool x_zero, double tol,
                          int maxit, std::vector<SolveStat>& out, int prec) {
  if (!ok()) return;
  if (!scratch_ || static_scratch_bytes(nrhs) > kStaticScratchBytes) {
    set_error("bicgs_static: static scratch unavailable or too small");
    return;
  }
  char* base = (char*)scratch_;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char* p = base + off;
    off += (bytes + 255) & ~(size_t)255;
    return p;
  };
  double2* part = (double2*)take((size_t)nrhs * kSlots * part_blocks_ * sizeof(double2));
  BiState* state = (BiState*)take((size_t)nrhs * sizeof(BiState));
  int* active = (int*)take((size_t)nrhs * sizeof(int));
  double* tolv = (double*)take((size_t)nrhs * sizeof(double));
  zd* lam = (zd*)take((size_t)nrhs * sizeof(zd));
  // the array keeps the previous call's contents: clear the scalars and flags
  const long words = (long)(((char*)tolv - (char*)state) / sizeof(unsigned));
  k_clear_words<<<nblocks(words), kBlock, 0, stream_>>>((unsigned*)state, words);
  ISG_HIP(hipGetLastError());
  double2* s_part = part_;
  BiState* s_state = state_;
  int* s_active = active_;
  double* s_tolv = tolv_;
  zd* s_lam = lam_;
  const bool s_graph = cfg_.graph, s_prof = prof_;
  part_ = part;
  state_ = state;
  active_ = active;
  tolv_ = tolv;
  lam_ = lam;
  cfg_.graph = false;
  prof_ = false;
  solve(x, g, gs, b, herm, nrhs, x_zero, tol, maxit, out, prec, false);
  part_ = s_part;
  state_ = s_state;
  active_ = s_active;
  tolv_ = s_tolv;
  lam_ = s_lam;
  cfg_.graph = s_graph;
  prof_ = s_prof;
}

// r_x = b - A x and A d with the solver's operator, lambda = <Ad, r_x> / <Ad, Ad>,
// x += lambda d, r0 = r_x - lambda Ad; BiCGSTAB then starts from (x, r0).
void Engine::bicgs_blend(zd* x, const zd* d, const zd* g, long gs, const zd* b, bool herm, int nrhs, double tol,
                         int maxit, std::vector<SolveStat>& out, int prec, std::vector<zd>* lam, bool x_zero,
                         std::vector<double>* ratio) {
  out.assign(nrhs, SolveStat{-1, 0, 0.0});
  if (!ok()) return;
  if (nrhs > cfg_.max_rhs) {
    set_error("bicgs: nrhs exceeds max_rhs");
    return;
  }
  const bool mixed = prec == PREC_MIXED && has32_;
  if (mixed)
    blend<float>(x, d, g, gs, b, herm, nrhs, PREC_MIXED, x_zero);
  else
    blend<double>(x, d, g, gs, b, herm, nrhs, PREC_FP64, x_zero);
  const int napp = x_zero ? 1 : 2;
  matvecs += napp * nrhs;
  if (mixed) matvecs32 += napp * nrhs;
  if (lam) {
    lam->resize(nrhs);
    ISG_HIP(hipMemcpyAsync(lam->data(), lam_, nrhs * sizeof(zd), hipMemcpyDeviceToHost, stream_));
    if (ratio) {
      ratio->resize(nrhs);
      ISG_HIP(hipMemcpyAsync(ratio->data(), bratio_, nrhs * sizeof(double), hipMemcpyDeviceToHost, stream_));
    }
    ISG_HIP(hipStreamSynchronize(stream_));
  }
  solve(x, g, gs, b, herm, nrhs, false, tol, maxit, out, prec, true);
}

template <typename V>
void Engine::blend(zd* x, const zd* d, const zd* g, long gs, const zd* b, bool herm, int nrhs, int prec,
                   bool x_zero) {
  const long N = P_.N;
  checked_x_ = nullptr;
  const int h = herm ? 1 : 0;
  dim3 grid(nblk_, nrhs), blk(kBlock);
  cplx<V>*y = vb<V>(V_Y), *op = vb<V>(V_OP), *res = vb<V>(V_RES), *ad = vb<V>(V_COR);
  if (!x_zero) {
    k_prep<V><<<grid, blk, 0, stream_>>>(x, g, gs, y, N, h);
    apply_vec<V>(y, op, nrhs, nullptr, prec);
  }
  k_bi_init<V><<<grid, blk, 0, stream_>>>(x, op, g, gs, b, nullptr, res, vb<V>(V_RT), N, h, x_zero ? 1 : 0, part_);
  k_prep<V><<<grid, blk, 0, stream_>>>(d, g, gs, y, N, h);
  apply_vec<V>(y, op, nrhs, nullptr, prec);
  k_bl_dots<V><<<grid, blk, 0, stream_>>>(d, op, g, gs, res, ad, N, h, part_);
  k_bl_fin<<<dim3(nrhs), blk, 0, stream_>>>(part_, nblk_, lam_, bratio_);
  k_bl_apply<V><<<grid, blk, 0, stream_>>>(x, d, res, ad, lam_, N);
}

void Engine::solve(zd* x, const zd* g, long gs, const zd* b, bool herm, int nrhs, bool x_zero, double tol,
                   int maxit, std::vector<SolveStat>& out, int prec, bool r0_in_res) {
  out.assign(nrhs, SolveStat{-1, 0, 0.0});
  if (!ok()) return;
  if (nrhs > cfg_.max_rhs) {
    set_error("bicgs: nrhs exceeds max_rhs");
    return;
  }
  const bool mixed = prec == PREC_MIXED && has32_;
  for (int i = 0; i < nrhs; i++) tol_h_[i] = mixed ? cfg_.mixed_margin * tol : tol;
  if (mixed)
    run_bicgs<float>(x, g, gs, b, herm, nrhs, x_zero, maxit, PREC_MIXED, r0_in_res ? vb<float>(V_RES) : nullptr);
  else
    run_bicgs<double>(x, g, gs, b, herm, nrhs, x_zero, maxit, PREC_FP64, r0_in_res ? vb<double>(V_RES) : nullptr);
  if (!ok()) return;
  for (int i = 0; i < nrhs; i++) {
    out[i].status = state_h_[i].status;
    out[i].iters = state_h_[i].iter;
    out[i].relres = state_h_[i].err;
  }
  if (mixed) {
    // FP64 true residual; defect correction with the FP32 operator until the
    // FP64 relative residual meets tol (or the rounds are used up)
    zd *res = vec_[V_RES], *cor = vec_[V_COR];
    std::vector<double>&rn = rn_, &bn = bn_;
    for (int round = 0; round <= kRefineRounds; round++) {
      true_residual(x, g, gs, b, herm, nrhs, rn, bn);
      if (!ok()) return;
      bool any = false;
      for (int i = 0; i < nrhs; i++) {
        double bnrm = bn[i] < 1e-20 ? 1.0 : bn[i];
        double err = rn[i] / bnrm;
        out[i].relres = err;
        const bool broke = out[i].status == 2 || out[i].status == 3;
        const bool need = !(err < tol) && !broke;
        if (err < tol)
          out[i].status = 0;
        else if (!broke)
          out[i].status = 1;
        // correction target relative to |res|, with the same margin; systems
        // that are done get an unreachable target and stay inactive
        tol_h_[i] = need ? cfg_.mixed_margin * tol * bnrm / std::max(rn[i], 1e-300) : 1e300;
        any = any || need;
      }
      if (!any || round == kRefineRounds) break;
      refinements++;
      // the residual vector itself, from the FP64 product the check left in V_OP
      k_bi_init<double><<<dim3(nblk_, nrhs), dim3(kBlock), 0, stream_>>>(x, vec_[V_OP], g, gs, b, nullptr, res,
                                                                        nullptr, P_.N, herm ? 1 : 0, 0, part_);
      ISG_HIP(hipMemsetAsync(cor, 0, (size_t)nrhs * P_.N * sizeof(zd), stream_));
      run_bicgs<float>(cor, g, gs, res, herm, nrhs, true, maxit, PREC_MIXED, nullptr);
      if (!ok()) return;
      k_add<<<dim3(nblk_, nrhs), dim3(kBlock), 0, stream_>>>(x, cor, P_.N);
      for (int i = 0; i < nrhs; i++) out[i].iters += state_h_[i].iter;
    }
    // every round ends with a check, so V_OP holds G prep(x) of the final x
    if (ok()) {
      checked_x_ = x;
      checked_n_ = nrhs;
      checked_h_ = herm;
    }
  }
  for (int i = 0; i < nrhs; i++) {
    solves++;
    solve_iters += out[i].iters;
  }
}

// ---------------------------------------------------------------- autotuner
template <typename T>
void Engine::autotune_prec(int nrhs, int reps) {
  const int L = P_.L, Lf = L - 1;
  Tuning& tu = tune(sizeof(T) == 4 ? PREC_MIXED : PREC_FP64, nrhs);
  zd *y = vec_[V_Y], *op = vec_[V_OP];
  checked_x_ = nullptr;
  ISG_HIP(hipMemsetAsync(y, 0, (size_t)nrhs * P_.N * sizeof(zd), stream_));
  apply<T, double>(y, op, nrhs, nullptr);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ISG_HIP(hipEventCreate(&e0));
  ISG_HIP(hipEventCreate(&e1));
  if (!ok()) return;
  // average time of one launch sequence (after one untimed run)
  auto timed = [&](auto&& run) -> float {
    run();
    ISG_HIP(hipEventRecord(e0, stream_));
    for (int r = 0; r < reps; r++) run();
    ISG_HIP(hipEventRecord(e1, stream_));
    ISG_HIP(hipEventSynchronize(e1));
    float ms = 0;
    ISG_HIP(hipEventElapsedTime(&ms, e0, e1));
    return ms / reps;
  };
  auto pick = [&](int ncand, auto&& admissible, auto&& run) -> int {
    int best = -1;
    float tbest = 0;
    for (int c = 0; c < ncand; c++) {
      if (!admissible(c)) continue;
      float t = timed([&] { run(c); });
      if (best < 0 || t < tbest) {
        best = c;
        tbest = t;
      }
    }
    return best;
  };
  // near field: tile widths, and for the FP32 operator the split-FP16 kernel,
  // timed on vectors of the operator's own type (those of the solver's products)
  const int widths[5] = {16, 32, 64, kNearSplit, kNearSplitWide};
  auto width_ok = [&](int c) {
    if (c == 3) return sizeof(T) == 4 && P_.B % 64 == 0 && near16_ != nullptr;
    if (c == 4) return sizeof(T) == 4 && P_.B % 128 == 0 && near16_ != nullptr;
    return P_.B % widths[c] == 0 && (widths[c] < 64 || sizeof(T) == 4);
  };
  int c = pick(5, width_ok, [&](int c) {
    launch_near<T, T>((const cplx<T>*)y, (cplx<T>*)op, nrhs, nullptr, true, widths[c]);
  });
  tu.near_nb = c < 0 ? -1 : widths[c];
  auto p2m_ok = [&](int c) {
    if (c == 3) return sizeof(T) == 4 && P_.B % 32 == 0 && p2m16_ != nullptr;
    return P_.B % widths[c] == 0 && (widths[c] < 64 || sizeof(T) == 4);
  };
  c = pick(4, p2m_ok, [&](int c) { launch_p2m<T, T>((const cplx<T>*)y, nrhs, nullptr, widths[c]); });
  tu.p2m_nb = c < 0 ? -1 : widths[c];
  for (int i = 2; i <= Lf; i++)
    tu.m2l[i] = pick(
        kM2LVariants, [&](int v) { return m2l_ok<T>(v, i); }, [&](int v) { launch_m2l<T>(i, nrhs, nullptr, v); });
  const std::vector<InterpShape>& sm = sizeof(T) == 4 ? m2m32_ : m2m64_;
  const std::vector<InterpShape>& sl = sizeof(T) == 4 ? l2l32_ : l2l64_;
  for (int i = 0; i + 1 < L; i++) {
    tu.m2m[i] = pick(
        5, [&](int k) { return kind_ok(sm[i], k); }, [&](int k) { launch_m2m<T>(i, nrhs, nullptr, k); });
    tu.l2l[i] = pick(
        5, [&](int k) { return kind_ok(sl[i], k); }, [&](int k) { launch_l2l<T>(i, true, nrhs, nullptr, k); });
  }
  // fused leaf kernel against the best separate leaf M2L + last L2L
  tu.leaf = -1;
  if (leaf_fuse_ok<T>()) {
    const int vm = tu.m2l[Lf], kl = tu.l2l[L - 2];
    tu.leaf = pick(
        2, [](int) { return true; },
        [&](int f) {
          if (f == 1) {
            launch_m2l_leaf<T>(nrhs, nullptr);
          } else {
            launch_m2l<T>(Lf, nrhs, nullptr, vm);
            launch_l2l<T>(L - 2, true, nrhs, nullptr, kl);
          }
        });
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  ISG_HIP(hipGetLastError());
}

void Engine::autotune(int nrhs, int reps) {
  if (!ok()) return;
  if (nrhs < 1 || nrhs > cfg_.max_rhs || reps < 1) {
    set_error("autotune: nrhs must be in 1..max_rhs and reps >= 1");
    return;
  }
  autotune_prec<double>(nrhs, reps);
  if (has32_) autotune_prec<float>(nrhs, reps);
}

static const char* kStageName[Engine::ST_COUNT] = {"near", "p2m", "m2l", "m2m", "l2l", "leaf"};

bool Engine::tuning_set(int prec, int cls, int stage, int level, int value) {
  if ((prec != PREC_FP64 && prec != PREC_MIXED) || cls < 0 || cls > 1 || stage < 0 || stage >= ST_COUNT) return false;
  Tuning& t = tune_[prec == PREC_MIXED ? 1 : 0][cls];
  if (stage == ST_NEAR || stage == ST_P2M) {
    const bool split = prec == PREC_MIXED && (value == kNearSplit || (stage == ST_NEAR && value == kNearSplitWide));
    if (value != -1 && value != 16 && value != 32 && value != 64 && !split) return false;
    (stage == ST_NEAR ? t.near_nb : t.p2m_nb) = value;
    return true;
  }
  if (stage == ST_LEAF) {
    if (value < -1 || value > 1) return false;
    t.leaf = value;
    return true;
  }
  if (level < 0 || level >= P_.L || value < -1) return false;
  if (stage == ST_M2L) {
    if (value >= kM2LVariants) return false;
    t.m2l[level] = value;
  } else {
    if (value > 4) return false;
    (stage == ST_M2M ? t.m2m : t.l2l)[level] = value;
  }
  return true;
}

int Engine::tuning_get(int prec, int cls, int stage, int level) const {
  if ((prec != PREC_FP64 && prec != PREC_MIXED) || cls < 0 || cls > 1 || stage < 0 || stage >= ST_COUNT) return -1;
  const Tuning& t = tune_[prec == PREC_MIXED ? 1 : 0][cls];
  if (stage == ST_NEAR) return t.near_nb;
  if (stage == ST_P2M) return t.p2m_nb;
  if (stage == ST_LEAF) return t.leaf;
  if (level < 0 || level >= P_.L) return -1;
  return stage == ST_M2L ? t.m2l[level] : (stage == ST_M2M ? t.m2m[level] : t.l2l[level]);
}

// Text cache, one choice per line: "depth precision class stage level value".
bool Engine::tuning_save(const std::string& path) {
  std::ofstream f(path);
  if (!f) return false;
  f << "# isg tuning: tree-depth precision(0 fp64, 1 mixed) batch-class(0 nrhs<16, 1 nrhs>=16) stage level value\n";
  for (int p = 0; p < 2; p++)
    for (int c = 0; c < 2; c++)
      for (int s = 0; s < ST_COUNT; s++)
        for (int i = 0; i < ((s == ST_NEAR || s == ST_P2M || s == ST_LEAF) ? 1 : P_.L); i++) {
          int v = tuning_get(p, c, s, i);
          if (v >= 0) f << P_.L << ' ' << p << ' ' << c << ' ' << kStageName[s] << ' ' << i << ' ' << v << '\n';
        }
  return (bool)f;
}

bool Engine::tuning_load(const std::string& path) {
  std::ifstream f(path);
  if (!f) return false;
  std::string line;
  while (std::getline(f, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream is(line);
    int depth, p, c, i, v;
    std::string name;
    if (!(is >> depth >> p >> c >> name >> i >> v)) return false;
    if (depth != P_.L) continue;
    int s = -1;
    for (int k = 0; k < ST_COUNT; k++)
      if (name == kStageName[k]) s = k;
    if (s < 0 || !tuning_set(p, c, s, i, v)) return false;
  }
  return true;
}

void Engine::stage(int which, int level, int rhs, std::vector<zd>& out) {
  out.clear();
  if (!ok() || level < 0 || level >= P_.L || rhs < 0 || rhs >= cfg_.max_rhs) return;
  const zd* base = (which == 0 ? op64_.agm : op64_.agl) + lvl_off_[level] + (size_t)rhs * lvl_size_[level];
  out.resize(lvl_size_[level]);
  ISG_HIP(hipStreamSynchronize(stream_));
  ISG_HIP(hipMemcpy(out.data(), base, out.size() * sizeof(zd), hipMemcpyDeviceToHost));
}

}  // namespace isg
