// isg_engine.hip -- device tables, MLFMA orchestration, receivers and the
// batched BiCGSTAB driver of the MI355X solver.
#include "isg_engine.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>

#include "isg_kernels.hpp"
#include "isg_solver_kernels.hpp"

namespace isg {

#define ISG_HIP(call)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess) {                                                                \
      set_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " #call);        \
    }                                                                                      \
  } while (0)

namespace {
inline unsigned nblocks(long work) { return (unsigned)((work + kBlock - 1) / kBlock); }
enum { V_R = 0, V_RT, V_P, V_V, V_S, V_T, V_Y, V_OP, V_AUX, V_COUNT };
}  // namespace

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
  d.near = up_c(P_.near);
  d.coeffT = up_c(P_.coeffT);
  d.basisT = up_c(P_.basisT);
  d.interpT.assign(L, nullptr);
  d.anterpT.assign(L, nullptr);
  d.shiftM.assign(L, nullptr);
  d.shiftL.assign(L, nullptr);
  d.trans.assign(L, nullptr);
  for (int i = 0; i < L; i++) {
    if (!P_.interpT[i].empty()) d.interpT[i] = up_r(P_.interpT[i]);
    if (!P_.anterpT[i].empty()) d.anterpT[i] = up_r(P_.anterpT[i]);
    if (!P_.shiftM[i].empty()) d.shiftM[i] = up_c(P_.shiftM[i]);
    if (!P_.shiftL[i].empty()) d.shiftL[i] = up_c(P_.shiftL[i]);
    if (!P_.trans[i].empty()) d.trans[i] = up_c(P_.trans[i]);
  }
  d.rxT = up_r(P_.rxT);
  d.rxaT = up_r(P_.rxaT);
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
  upload(op64_);
  auto up_i = [&](const std::vector<int>& src) -> int* {
    int* p = (int*)dalloc(src.size() * sizeof(int));
    if (p) ISG_HIP(hipMemcpy(p, src.data(), src.size() * sizeof(int), hipMemcpyHostToDevice));
    return p;
  };
  int L = P_.L, R = cfg_.max_rhs;
  interp0_.assign(L, nullptr);
  anterp0_.assign(L, nullptr);
  for (int i = 0; i < L; i++) {
    if (!P_.interp0[i].empty()) interp0_[i] = up_i(P_.interp0[i]);
    if (!P_.anterp0[i].empty()) anterp0_[i] = up_i(P_.anterp0[i]);
  }
  rx0_ = up_i(P_.rx0);
  rxaI_ = up_i(P_.rxaI);
  // multipole and local stores, all levels, [level][rhs][4^i S_i]
  lvl_off_.assign(L, 0);
  lvl_size_.assign(L, 0);
  size_t tot = 0;
  for (int i = 0; i < L; i++) {
    lvl_size_[i] = (size_t)P_.nclus[i] * P_.nsamp[i];
    lvl_off_[i] = tot;
    tot += lvl_size_[i] * R;
  }
  agm_ = (zd*)dalloc(tot * sizeof(zd));
  agl_ = (zd*)dalloc(tot * sizeof(zd));
  vec_.assign(V_COUNT, nullptr);
  for (int k = 0; k < V_COUNT; k++) vec_[k] = (zd*)dalloc((size_t)R * P_.N * sizeof(zd));
  nblk_ = (int)std::min<long>(1024, std::max<long>(1, (P_.N + kBlock * 4 - 1) / (kBlock * 4)));
  part_ = (double2*)dalloc((size_t)R * kSlots * nblk_ * sizeof(double2));
  state_ = (BiState*)dalloc((size_t)R * sizeof(BiState));
  active_ = (int*)dalloc((size_t)R * sizeof(int));
  ISG_HIP(hipHostMalloc((void**)&active_h_, (size_t)R * sizeof(int), hipHostMallocDefault));
  ISG_HIP(hipHostMalloc((void**)&state_h_, (size_t)R * sizeof(BiState), hipHostMallocDefault));
  if (ok()) {
    ISG_HIP(hipMemset(state_, 0, (size_t)R * sizeof(BiState)));
    ISG_HIP(hipMemset(active_, 0, (size_t)R * sizeof(int)));
  }
}

Engine::~Engine() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (auto& k : kt_) {
    for (auto e : k.ev0) (void)hipEventDestroy(e);
    for (auto e : k.ev1) (void)hipEventDestroy(e);
  }
  for (void* p : allocs_) (void)hipFree(p);
  if (active_h_) (void)hipHostFree(active_h_);
  if (state_h_) (void)hipHostFree(state_h_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

// r = G y (y already contrast-weighted).  Stage order of matvec_gpu.cu:103-264:
// P2M, M2M up to level 2, M2L on levels 2..leaf, L2L down, L2P, then P2P.
void Engine::apply_mlfma(const zd* y, zd* r, int nrhs, const int* act) {
  if (!ok()) return;
  const int L = P_.L, Lf = L - 1, N = P_.N, B = P_.B;
  const DevOp<double>& d = op64_;
  const int S = P_.nsamp[Lf];
  dim3 blk(kBlock);
  auto M = [&](int i) { return agm_ + lvl_off_[i]; };
  auto Lc = [&](int i) { return agl_ + lvl_off_[i]; };
  auto sz = [&](int i) { return (long)lvl_size_[i]; };
  k_p2m<double, double><<<dim3(nblocks((long)B * S), nrhs), blk, 0, stream_>>>(y, N, M(Lf), sz(Lf), d.coeffT, S, B, act);
  for (int i = L - 2; i >= 2; i--)
    k_m2m<double><<<dim3(nblocks(sz(i)), nrhs), blk, 0, stream_>>>(M(i + 1), sz(i + 1), M(i), sz(i), d.interpT[i],
                                                                      interp0_[i], d.shiftM[i], P_.nsamp[i],
                                                                      P_.nsamp[i + 1], P_.nclus[i], cfg_.plan.ninter, act);
  prof_begin(KT_M2L);
  for (int i = 2; i <= Lf; i++)
    k_m2l<double><<<dim3(nblocks(sz(i)), nrhs), blk, 0, stream_>>>(M(i), Lc(i), sz(i), d.trans[i], i, P_.nsamp[i], act);
  prof_end(KT_M2L);
  for (int i = 2; i <= L - 2; i++)
    k_l2l<double><<<dim3(nblocks(sz(i + 1)), nrhs), blk, 0, stream_>>>(Lc(i), sz(i), Lc(i + 1), sz(i + 1), d.anterpT[i + 1],
                                                                          anterp0_[i + 1], d.shiftL[i], P_.nsamp[i],
                                                                          P_.nsamp[i + 1], P_.nclus[i],
                                                                          2 * cfg_.plan.ninter, 1, act);
  launch_near(y, Lc(Lf), r, nrhs, act, true);
  ISG_HIP(hipGetLastError());
}

// Near field + L2P on the matrix cores (k_near_mfma), or L2P alone when
// with_near is false (receiver back-projection).
void Engine::launch_near(const zd* y, const zd* loc, zd* r, int nrhs, const int* act, bool with_near) {
  const int Lf = P_.L - 1, S = P_.nsamp[Lf], B = P_.B;
  const long N = P_.N, ls = (long)lvl_size_[Lf];
  const int nb = B >= 32 ? 32 : 16;
  size_t lds = (size_t)nb * near_row(S) * sizeof(zd);
  dim3 grid(B / nb, nrhs), blk(kBlock);
  if (with_near) prof_begin(KT_NEAR);
  if (nb == 32)
    k_near_mfma<double, double, 32><<<grid, blk, lds, stream_>>>(y, N, loc, ls, r, N, op64_.near, op64_.basisT, S,
                                                                 1 << Lf, with_near ? 1 : 0, act);
  else
    k_near_mfma<double, double, 16><<<grid, blk, lds, stream_>>>(y, N, loc, ls, r, N, op64_.near, op64_.basisT, S,
                                                                 1 << Lf, with_near ? 1 : 0, act);
  if (with_near) prof_end(KT_NEAR);
}

// ---- live kernel timing (HIP events on the engine stream)
void Engine::profile(bool on) {
  harvest();
  prof_ = on;
  for (auto& k : kt_) {
    k.ms = 0;
    k.launches = 0;
  }
}

void Engine::prof_begin(int id) {
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
    }
  }
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
        k.launches++;
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

void Engine::mlfma(const zd* x, zd* r, int nrhs, const int* act) {
  apply_mlfma(x, r, nrhs, act);
  matvecs += nrhs;
}

// matvec_gpu.cu:266-368 with the top levels and the receiver interpolation
// on the device as well.
void Engine::aggregate(const zd* x, zd* r, int nrhs) {
  if (!ok()) return;
  const int L = P_.L, Lf = L - 1, N = P_.N, B = P_.B, S = P_.nsamp[Lf];
  const DevOp<double>& d = op64_;
  auto M = [&](int i) { return agm_ + lvl_off_[i]; };
  auto sz = [&](int i) { return (long)lvl_size_[i]; };
  dim3 blk(kBlock);
  k_p2m<double, double><<<dim3(nblocks((long)B * S), nrhs), blk, 0, stream_>>>(x, N, M(Lf), sz(Lf), d.coeffT, S, B, nullptr);
  for (int i = L - 2; i >= 0; i--)
    k_m2m<double><<<dim3(nblocks(sz(i)), nrhs), blk, 0, stream_>>>(M(i + 1), sz(i + 1), M(i), sz(i), d.interpT[i],
                                                                      interp0_[i], d.shiftM[i], P_.nsamp[i],
                                                                      P_.nsamp[i + 1], P_.nclus[i], cfg_.plan.ninter,
                                                                      nullptr);
  int R = cfg_.plan.numrx;
  k_rx_interp<double><<<dim3(nblocks(R), nrhs), blk, 0, stream_>>>(M(0), sz(0), r, R, d.rxT, rx0_, P_.nsamp[0], R,
                                                                    cfg_.plan.ninter);
  ISG_HIP(hipGetLastError());
}

// matvec_gpu.cu:369-470 (argument order out x, in r as in the reference),
// every level on the device.
void Engine::aggregateh(const zd* r, zd* x, int nrhs) {
  if (!ok()) return;
  const int L = P_.L, Lf = L - 1, N = P_.N, B = P_.B, S = P_.nsamp[Lf];
  const DevOp<double>& d = op64_;
  auto Lc = [&](int i) { return agl_ + lvl_off_[i]; };
  auto sz = [&](int i) { return (long)lvl_size_[i]; };
  dim3 blk(kBlock);
  int R = cfg_.plan.numrx;
  double sc = (double)R * cfg_.plan.res * cfg_.plan.res;
  k_rx_anterp<double><<<dim3(nblocks(P_.nsamp[0]), nrhs), blk, 0, stream_>>>(r, R, Lc(0), sz(0), d.rxaT, rxaI_,
                                                                              P_.nsamp[0], 2 * cfg_.plan.ninter, sc);
  for (int i = 0; i <= L - 2; i++)
    k_l2l<double><<<dim3(nblocks(sz(i + 1)), nrhs), blk, 0, stream_>>>(Lc(i), sz(i), Lc(i + 1), sz(i + 1), d.anterpT[i + 1],
                                                                          anterp0_[i + 1], d.shiftL[i], P_.nsamp[i],
                                                                          P_.nsamp[i + 1], P_.nclus[i],
                                                                          2 * cfg_.plan.ninter, 0, nullptr);
  launch_near(nullptr, Lc(Lf), x, nrhs, nullptr, false);
  ISG_HIP(hipGetLastError());
}

void Engine::launch_prep(const zd* x, const zd* g, long gs, zd* y, bool herm, int nrhs) {
  k_prep<<<dim3(nblk_, nrhs), dim3(kBlock), 0, stream_>>>(x, g, gs, y, P_.N, herm ? 1 : 0);
}

void Engine::reduce_finalize(int phase, int nrhs, double tol, int maxit) {
  k_bi_fin<<<dim3(nrhs), dim3(kBlock), 0, stream_>>>(phase, state_, active_, part_, nblk_, tol, maxit);
}

// bicgs_gpu.cu:116-224, batched over right-hand sides with device-resident
// scalars; the host reads only the per-right-hand-side activity flags once
// per iteration to know when every system has stopped.
void Engine::bicgs(zd* x, const zd* g, long gs, const zd* b, bool herm, int nrhs, bool x_zero, double tol,
                   int maxit, std::vector<SolveStat>& out) {
  out.assign(nrhs, SolveStat{-1, 0, 0.0});
  if (!ok()) return;
  if (nrhs > cfg_.max_rhs) {
    set_error("bicgs: nrhs exceeds max_rhs");
    return;
  }
  const long N = P_.N;
  zd *r = vec_[V_R], *rt = vec_[V_RT], *p = vec_[V_P], *v = vec_[V_V], *s = vec_[V_S], *t = vec_[V_T],
     *y = vec_[V_Y], *op = vec_[V_OP];
  const int h = herm ? 1 : 0;
  dim3 grid(nblk_, nrhs), blk(kBlock);
  if (!x_zero) {
    k_prep<<<grid, blk, 0, stream_>>>(x, g, gs, y, N, h);
    apply_mlfma(y, op, nrhs, nullptr);
    matvecs += nrhs;
  }
  k_bi_init<<<grid, blk, 0, stream_>>>(x, op, g, gs, b, r, rt, N, h, x_zero ? 1 : 0, part_);
  reduce_finalize(PH_INIT, nrhs, tol, maxit);
  for (int it = 0; it <= maxit; it++) {
    ISG_HIP(hipMemcpyAsync(active_h_, active_, nrhs * sizeof(int), hipMemcpyDeviceToHost, stream_));
    ISG_HIP(hipStreamSynchronize(stream_));
    if (!ok()) return;
    if (prof_) harvest();
    int nact = 0;
    for (int i = 0; i < nrhs; i++) nact += active_h_[i] != 0;
    if (nact == 0) break;
    k_bi_pupd<<<grid, blk, 0, stream_>>>(r, p, v, g, gs, y, N, h, state_);
    apply_mlfma(y, op, nrhs, active_);
    k_bi_post1<<<grid, blk, 0, stream_>>>(p, op, g, gs, v, rt, N, h, state_, part_);
    reduce_finalize(PH_ALPHA, nrhs, tol, maxit);
    k_bi_supd<<<grid, blk, 0, stream_>>>(r, v, g, gs, s, y, N, h, state_);
    apply_mlfma(y, op, nrhs, active_);
    k_bi_post2<<<grid, blk, 0, stream_>>>(s, op, g, gs, t, N, h, state_, part_);
    reduce_finalize(PH_OMEGA, nrhs, tol, maxit);
    prof_begin(KT_VEC);
    k_bi_xrupd<<<grid, blk, 0, stream_>>>(x, p, s, t, r, rt, N, state_, part_);
    prof_end(KT_VEC);
    reduce_finalize(PH_RESID, nrhs, tol, maxit);
    matvecs += 2 * nact;
  }
  ISG_HIP(hipMemcpyAsync(state_h_, state_, nrhs * sizeof(BiState), hipMemcpyDeviceToHost, stream_));
  ISG_HIP(hipStreamSynchronize(stream_));
  for (int i = 0; i < nrhs; i++) {
    out[i].status = state_h_[i].status;
    out[i].iters = state_h_[i].iter;
    out[i].relres = state_h_[i].err;
    solves++;
    solve_iters += state_h_[i].iter;
  }
}

void Engine::stage(int which, int level, int rhs, std::vector<zd>& out) {
  out.clear();
  if (!ok() || level < 0 || level >= P_.L || rhs < 0 || rhs >= cfg_.max_rhs) return;
  const zd* base = (which == 0 ? agm_ : agl_) + lvl_off_[level] + (size_t)rhs * lvl_size_[level];
  out.resize(lvl_size_[level]);
  ISG_HIP(hipStreamSynchronize(stream_));
  ISG_HIP(hipMemcpy(out.data(), base, out.size() * sizeof(zd), hipMemcpyDeviceToHost));
}

}  // namespace isg
