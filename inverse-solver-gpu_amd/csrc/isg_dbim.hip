// isg_dbim.hip -- DBIM driver: vector algebra kernels and the iteration.
#include "isg_dbim.hpp"

#include <chrono>
#include <cmath>
#include <cstring>

#include "isg_kernels.hpp"

namespace isg {

#define DB_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess && err_.empty()) err_ = std::string("HIP: ") + hipGetErrorString(e_) + " at " #call; \
  } while (0)

#define DB_LOOP(N) for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < (N); i += (long)gridDim.x * kBlock)

namespace {

// inc[t][n] = exp(i k0 u_t . r_n): plane wave along transmitter t
// (setup_born.cpp:62-65); with radius > 0 the field of a line source at
// radius * u_t instead: (i/4) H0^(1)(k0 |r_n - radius u_t|).
__global__ void __launch_bounds__(kBlock) k_incident(zd* inc, const zd* pos, const zd* tx, long N, double k0,
                                                     double radius) {
  int t = blockIdx.y;
  zd u = tx[t];
  DB_LOOP(N) {
    zd p = pos[i];
    if (radius > 0) {
      const double dx = p.x - radius * u.x, dy = p.y - radius * u.y;
      const double kd = k0 * sqrt(dx * dx + dy * dy);
      inc[t * N + i] = mkc<double>(-0.25 * y0(kd), 0.25 * j0(kd));
    } else {
      double ph = k0 * (u.x * p.x + u.y * p.y);
      double s, c;
      sincos(ph, &s, &c);
      inc[t * N + i] = mkc<double>(c, s);
    }
  }
}

// g = o k0 k0
__global__ void __launch_bounds__(kBlock) k_weight(const zd* o, zd* g, long N, double k0) {
  DB_LOOP(N) {
    zd v = o[i];
    g[i] = mkc<double>(v.x * k0 * k0, v.y * k0 * k0);
  }
}

// out[t] = a[t] * w  (w shared)
__global__ void __launch_bounds__(kBlock) k_rowmul(const zd* a, const zd* w, zd* out, long N) {
  int t = blockIdx.y;
  DB_LOOP(N) out[t * N + i] = cmul(a[t * N + i], w[i]);
}

// y[t] = tot[t] * ot * k0^2   (mom.cpp:284)
__global__ void __launch_bounds__(kBlock) k_deriv_src(const zd* tot, const zd* ot, zd* y, long N, double k0) {
  int t = blockIdx.y;
  DB_LOOP(N) {
    zd v = cmul(tot[t * N + i], ot[i]);
    y[t * N + i] = mkc<double>(v.x * k0 * k0, v.y * k0 * k0);
  }
}

// y[t] += g * x[t]   (mom.cpp:301)
__global__ void __launch_bounds__(kBlock) k_add_gx(zd* y, const zd* g, const zd* x, long N) {
  int t = blockIdx.y;
  DB_LOOP(N) y[t * N + i] = cadd(y[t * N + i], cmul(g[i], x[t * N + i]));
}

// c[t] = conj(g) * y[t]   (mom.cpp:344)
__global__ void __launch_bounds__(kBlock) k_adj_src(const zd* g, const zd* y, zd* c, long N) {
  int t = blockIdx.y;
  DB_LOOP(N) c[t * N + i] = cmul(cconj(g[i]), y[t * N + i]);
}

// y = conj(x)
__global__ void __launch_bounds__(kBlock) k_conj(const zd* x, zd* y, long N) {
  int t = blockIdx.y;
  DB_LOOP(N) y[t * N + i] = cconj(x[t * N + i]);
}

// gradient accumulation over a batch, transmitters in order (mom.cpp:351-360):
// for t: y_t += conj(Gc_t); buff += conj(tot_t) y_t; buff += reg^2 o
__global__ void __launch_bounds__(kBlock) k_grad_acc(zd* buff, const zd* tot, const zd* y, const zd* gcx, const zd* o,
                                                     long N, int nb, double reg2) {
  DB_LOOP(N) {
    zd acc = buff[i];
    zd ov = o[i];
    for (int t = 0; t < nb; t++) {
      zd yv = cadd(y[t * N + i], cconj(gcx[t * N + i]));
      acc = cadd(acc, cmul(cconj(tot[t * N + i]), yv));
      acc = cadd(acc, mkc<double>(reg2 * ov.x, reg2 * ov.y));
    }
    buff[i] = acc;
  }
}

// b = -a
__global__ void __launch_bounds__(kBlock) k_neg(const zd* a, zd* b, long N) {
  DB_LOOP(N) b[i] = mkc<double>(-a[i].x, -a[i].y);
}

// a = a + s * b  (complex s)
__global__ void __launch_bounds__(kBlock) k_axpy(zd* a, const zd* b, zd s, long N) {
  DB_LOOP(N) a[i] = cadd(a[i], cmul(s, b[i]));
}

// a[t] += s b[t] for every row t of a batch
__global__ void __launch_bounds__(kBlock) k_axpy_rows(zd* a, const zd* b, zd s, long N) {
  int t = blockIdx.y;
  DB_LOOP(N) a[t * N + i] = cadd(a[t * N + i], cmul(s, b[t * N + i]));
}

// Network input of transmitters t0..t0+nc-1, written straight into the
// network's space-to-depth layout: position (Y, X) of an image holds channel
// 8 s + c = raster pixel (2Y + s / 2, 2X + s % 2), channel c of
// (o, inc_t, tot_t, -alpha dtot_t) -- the field channels divided by the
// transmitter's incident-field rms (inv[t]) -- times the network's input
// scale sc[c]; tot_t is the first-order start tot_prev - alpha dtot.  One
// thread per (position, sub-pixel).
struct NetScales {
  float in[WarmNet::kInCh];
};
__global__ void __launch_bounds__(kBlock) k_net_in(const zd* o, const zd* inc, const zd* tot, const zd* dtot,
                                                   double alpha, const double* inv, const int* r2t, int side,
                                                   NetScales sc, float* x32) {
  const int t = blockIdx.y;
  const long N = (long)side * side;
  const double s = inv[t];
  const int hs = side / 2;
  DB_LOOP(N) {
    const int sp = (int)(i & 3);
    const long pos = i >> 2;
    const int X = (int)(pos % hs), Y = (int)(pos / hs);
    const long q = (long)(2 * Y + (sp >> 1)) * side + 2 * X + (sp & 1);
    const int n = r2t[q];
    const zd a = o[n], b = inc[t * N + n], c = tot[t * N + n], d = dtot[t * N + n];
    float* p = x32 + (t * (N >> 2) + pos) * 32 + 8 * sp;
    p[0] = (float)a.x * sc.in[0];
    p[1] = (float)a.y * sc.in[1];
    p[2] = (float)(b.x * s) * sc.in[2];
    p[3] = (float)(b.y * s) * sc.in[3];
    p[4] = (float)(c.x * s) * sc.in[4];
    p[5] = (float)(c.y * s) * sc.in[5];
    p[6] = (float)(-alpha * d.x * s) * sc.in[6];
    p[7] = (float)(-alpha * d.y * s) * sc.in[7];
  }
}

// d[t][r2t[q]] = the network's output for raster pixel q (channels 2 s, 2 s + 1
// of its position in the packed output, s its sub-pixel) * scale[t]
__global__ void __launch_bounds__(kBlock) k_net_dir(const float* y16, const double* scale, const int* r2t, int side,
                                                    zd* d) {
  const int t = blockIdx.y;
  const long N = (long)side * side;
  const double s = scale[t];
  const int hs = side / 2;
  DB_LOOP(N) {
    const int y = (int)(i / side), x = (int)(i % side);
    const int sp = 2 * (y & 1) + (x & 1);
    const float* p = y16 + (t * (N >> 2) + (long)(y >> 1) * hs + (x >> 1)) * 16 + 2 * sp;
    d[t * N + r2t[i]] = mkc<double>(p[0] * s, p[1] * s);
  }
}

// out[t] = a[t] * f, t < n
__global__ void k_scale_rows(const double* a, double f, int n, double* out) {
  for (int t = threadIdx.x; t < n; t += blockDim.x) out[t] = a[t] * f;
}

// part[t][block] = partial sum of |a[t][n]|^2 over the block's share of row t
__global__ void __launch_bounds__(kBlock) k_row_sq(const zd* a, long N, double* part) {
  const int t = blockIdx.y;
  double v = 0;
  DB_LOOP(N) {
    const zd x = a[t * N + i];
    v += x.x * x.x + x.y * x.y;
  }
  __shared__ double sh[kBlock / 64];
  const double w = wave_sum(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = w;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0;
    for (int j = 0; j < kBlock / 64; j++) s += sh[j];
    part[(long)t * gridDim.x + blockIdx.x] = s;
  }
}

// ot = od + beta ot
__global__ void __launch_bounds__(kBlock) k_newdir(zd* ot, const zd* od, zd beta, long N) {
  DB_LOOP(N) ot[i] = cadd(od[i], cmul(beta, ot[i]));
}

// partial sums for the metrics and the Polak-Ribiere coefficient:
// 0 |od|^2, 1-2 od conj(od - save), 3 |save|^2, 4 |o|^2, 5 |o - omes|^2
constexpr int kSums = 6;
__global__ void __launch_bounds__(kBlock) k_sums(const zd* o, const zd* omes, const zd* od, const zd* save, long N,
                                                 double* part) {
  double v[kSums] = {0, 0, 0, 0, 0, 0};
  DB_LOOP(N) {
    zd a = od[i], s = save[i], ov = o[i], om = omes[i];
    v[0] += a.x * a.x + a.y * a.y;
    zd d = csub(a, s);
    v[1] += a.x * d.x + a.y * d.y;
    v[2] += a.y * d.x - a.x * d.y;
    v[3] += s.x * s.x + s.y * s.y;
    v[4] += ov.x * ov.x + ov.y * ov.y;
    zd e = csub(ov, om);
    v[5] += e.x * e.x + e.y * e.y;
  }
  __shared__ double sh[kSums][kBlock / 64];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int s = 0; s < kSums; s++) {
    double a = wave_sum(v[s]);
    if (lane == 0) sh[s][w] = a;
  }
  __syncthreads();
  if (threadIdx.x < kSums) {
    double a = 0;
    for (int j = 0; j < kBlock / 64; j++) a += sh[threadIdx.x][j];
    part[threadIdx.x * gridDim.x + blockIdx.x] = a;
  }
}

}  // namespace

void* Dbim::dalloc(size_t bytes) {
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes ? bytes : 16);
  if (e != hipSuccess) {
    if (err_.empty()) err_ = std::string("hipMalloc: ") + hipGetErrorString(e);
    return nullptr;
  }
  allocs_.push_back(p);
  return p;
}

Dbim::Dbim(Engine& eng, const DbimOptions& opt) : eng_(eng), opt_(opt) {
  if (!eng_.ok()) return;
  const Plan& P = eng_.plan();
  N_ = P.N;
  R_ = P.cfg.numrx;
  T_ = opt_.numtx;
  B_ = eng_.max_rhs();
  k0_ = P.k0;
  if (T_ < 1) {
    err_ = "numtx must be >= 1";
    return;
  }
  size_t vb = sizeof(zd) * (size_t)N_;
  pos_ = (zd*)dalloc(vb);
  tx_ = (zd*)dalloc(sizeof(zd) * T_);
  o_ = (zd*)dalloc(vb);
  od_ = (zd*)dalloc(vb);
  ot_ = (zd*)dalloc(vb);
  save_ = (zd*)dalloc(vb);
  omes_ = (zd*)dalloc(vb);
  g_ = (zd*)dalloc(vb);
  buff_ = (zd*)dalloc(vb);
  od0_ = (zd*)dalloc(vb);
  inc_ = (zd*)dalloc(vb * T_);
  tot_ = (zd*)dalloc(vb * T_);
  mes_ = (zd*)dalloc(sizeof(zd) * (size_t)T_ * R_);
  mesd_ = (zd*)dalloc(sizeof(zd) * (size_t)T_ * R_);
  rhs_ = (zd*)dalloc(sizeof(zd) * (size_t)T_ * R_);
  wy_ = (zd*)dalloc(vb * B_);
  wc_ = (zd*)dalloc(vb * B_);
  wx_ = (zd*)dalloc(vb * B_);
  if (P.cfg.radius > 0 && B_ < T_) {
    if (eng_.rx_rhs() < T_) {
      err_ = "receiver operator capacity below numtx";
      return;
    }
    rxw_ = (zd*)dalloc(vb * T_);
  }
  if (opt_.field_start >= 2 || opt_.solve_start == 1) dtot_ = (zd*)dalloc(vb * T_);
  if (opt_.solve_start == 1) adj_ = (zd*)dalloc(vb * T_);
  nblk_ = (int)std::min<long>(1024, std::max<long>(1, (N_ + kBlock * 4 - 1) / (kBlock * 4)));
  part_ = (double*)dalloc(sizeof(double) * kSums * nblk_);
  DB_HIP(hipHostMalloc((void**)&part_h_, sizeof(double) * kSums * nblk_, hipHostMallocDefault));
  if (!ok()) return;
  DB_HIP(hipMemcpy(pos_, P.pos.data(), vb, hipMemcpyHostToDevice));
  // transmitter directions equispaced on 360 degrees (mom.cpp:141-147)
  std::vector<zd> tx(T_);
  for (int i = 0; i < T_; i++) {
    double ang = 360.0 / T_ * i;
    tx[i] = mkc<double>(std::cos(ang / 360 * 2 * M_PI), std::sin(ang / 360 * 2 * M_PI));
  }
  DB_HIP(hipMemcpy(tx_, tx.data(), sizeof(zd) * T_, hipMemcpyHostToDevice));
}

Dbim::~Dbim() {
  (void)hipStreamSynchronize(eng_.stream());
  for (void* p : allocs_) (void)hipFree(p);
  if (part_h_) (void)hipHostFree(part_h_);
  if (rowpart_h_) (void)hipHostFree(rowpart_h_);
  if (step_rms_h_) (void)hipHostFree(step_rms_h_);
}

void Dbim::solve(zd* x, const zd* b, bool herm, int nb, bool x_zero, DbimStats& st) {
  std::vector<SolveStat> ss;
  eng_.bicgs(x, g_, 0, b, herm, nb, x_zero, opt_.tol, opt_.maxit, ss, opt_.precision);
  for (auto& s : ss) {
    st.solves++;
    st.solve_iters += s.iters;
    if (s.status != 0) st.not_converged++;
  }
}

void Dbim::solve_from(zd* x, zd* prev, const zd* b, bool herm, int nb, DbimStats& st) {
  hipStream_t s = eng_.stream();
  const size_t bytes = sizeof(zd) * (size_t)N_ * nb;
  DB_HIP(hipMemsetAsync(x, 0, bytes, s));
  std::vector<SolveStat> ss;
  eng_.bicgs_blend(x, prev, g_, 0, b, herm, nb, opt_.tol, opt_.maxit, ss, opt_.precision, nullptr, true);
  for (auto& r : ss) {
    st.solves++;
    st.solve_iters += r.iters;
    if (r.status != 0) st.not_converged++;
  }
  // the copy leaves the engine's operator output (a mixed solve's checked product) untouched
  DB_HIP(hipMemcpyAsync(prev, x, bytes, hipMemcpyDeviceToDevice, s));
}

// mom.cpp:247-271 / :338-362: od = sum_t conj(tot_t) (A_t^H mesd_t) + reg^2 o
void Dbim::gradient(DbimStats& st) {
  hipStream_t s = eng_.stream();
  const long N = N_;
  DB_HIP(hipMemsetAsync(buff_, 0, sizeof(zd) * N, s));
  // the receiver back-projection of every transmitter in one call when the
  // transmitters run in several batches (the dense finite-radius operator is
  // then read once per phase)
  if (rxw_) eng_.aggregateh(mesd_, rxw_, T_);
  for (int t0 = 0; t0 < T_; t0 += B_) {
    int nb = std::min(B_, T_ - t0);
    dim3 grid(nblk_, nb), blk(kBlock);
    zd* wy = rxw_ ? rxw_ + (size_t)t0 * N : wy_;
    if (!rxw_) eng_.aggregateh(mesd_ + (size_t)t0 * R_, wy, nb);
    k_adj_src<<<grid, blk, 0, s>>>(g_, wy, wc_, N);
    if (adj_ && adj_ok_) {
      solve_from(wx_, adj_ + (size_t)t0 * N, wc_, true, nb, st);
    } else {
      DB_HIP(hipMemsetAsync(wx_, 0, sizeof(zd) * N * nb, s));
      solve(wx_, wc_, true, nb, true, st);
      if (adj_) DB_HIP(hipMemcpyAsync(adj_ + (size_t)t0 * N, wx_, sizeof(zd) * N * nb, hipMemcpyDeviceToDevice, s));
    }
    // c = conj(G conj(x))  (mlfmah, solve.cpp:8-24); the conjugate of Gc is taken in k_grad_acc.
    // A mixed-precision solve ends with an FP64 residual check that computed G conj(x) of the
    // final x: that is this product, with the FP64 operator, at no cost.  After an FP64 solve
    // it is computed here, also with the FP64 operator.
    const zd* gcx = eng_.checked_product(wx_, nb, true);
    if (!gcx) {
      k_conj<<<grid, blk, 0, s>>>(wx_, wc_, N);
      eng_.mlfma(wc_, wx_, nb, nullptr, PREC_FP64);
      gcx = wx_;
    }
    k_grad_acc<<<dim3(nblk_), blk, 0, s>>>(buff_, tot_ + (size_t)t0 * N, wy, gcx, o_, N, nb,
                                           opt_.regular * opt_.regular);
  }
  adj_ok_ = adj_ != nullptr;
  DB_HIP(hipMemcpyAsync(od_, buff_, sizeof(zd) * N, hipMemcpyDeviceToDevice, s));
}

// setup_born.cpp:171-242 metrics (hist.txt columns); all squared ratios
void Dbim::metrics(DbimStats& st) {
  hipStream_t s = eng_.stream();
  k_sums<<<dim3(nblk_), dim3(kBlock), 0, s>>>(o_, omes_, od_, save_, N_, part_);
  DB_HIP(hipMemcpyAsync(part_h_, part_, sizeof(double) * kSums * nblk_, hipMemcpyDeviceToHost, s));
  std::vector<zd> md((size_t)T_ * R_);
  DB_HIP(hipMemcpyAsync(md.data(), mesd_, sizeof(zd) * md.size(), hipMemcpyDeviceToHost, s));
  DB_HIP(hipStreamSynchronize(s));
  double sum[kSums] = {0, 0, 0, 0, 0, 0};
  for (int k = 0; k < kSums; k++)
    for (int j = 0; j < nblk_; j++) sum[k] += part_h_[k * nblk_ + j];
  double err = 0;
  for (auto& v : md) err += v.x * v.x + v.y * v.y;
  double reg = opt_.regular;
  st.hist[0] = err / mesnorm_;
  st.hist[1] = sum[5] / objnorm_;
  st.hist[2] = reg * reg * sum[4] / mesnorm_;
  st.hist[3] = sum[0];
  st.hist[4] = sum[4];
}

void Dbim::setup(const zd* otrue, DbimStats& st) {
  if (!ok()) return;
  auto t_start = std::chrono::steady_clock::now();
  hipStream_t s = eng_.stream();
  const long N = N_;
  dim3 blk(kBlock);
  k_incident<<<dim3(nblk_, T_), blk, 0, s>>>(inc_, pos_, tx_, N, k0_, eng_.plan().cfg.radius);
  DB_HIP(hipMemcpyAsync(o_, otrue, sizeof(zd) * N, hipMemcpyHostToDevice, s));
  k_weight<<<dim3(nblk_), blk, 0, s>>>(o_, g_, N, k0_);
  // synthetic measurements: mes_t = aggregate(g * A^-1 inc_t)  (setup_born.cpp:139-149)
  for (int t0 = 0; t0 < T_; t0 += B_) {
    int nb = std::min(B_, T_ - t0);
    DB_HIP(hipMemsetAsync(wx_, 0, sizeof(zd) * N * nb, s));
    solve(wx_, inc_ + (size_t)t0 * N, false, nb, true, st);
    zd* wy = rxw_ ? rxw_ + (size_t)t0 * N : wy_;
    k_rowmul<<<dim3(nblk_, nb), blk, 0, s>>>(wx_, g_, wy, N);
    if (!rxw_) eng_.aggregate(wy, mes_ + (size_t)t0 * R_, nb);
    if (progress_) {
      DB_HIP(hipStreamSynchronize(s));
      if (!ok()) return;
      progress_(t0 + nb, T_);
    }
  }
  if (rxw_) eng_.aggregate(rxw_, mes_, T_);
  std::vector<zd> mes((size_t)T_ * R_);
  DB_HIP(hipMemcpyAsync(mes.data(), mes_, sizeof(zd) * mes.size(), hipMemcpyDeviceToHost, s));
  DB_HIP(hipStreamSynchronize(s));
  start(mes.data(), otrue, st);
  st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
}

void Dbim::setup_measured(const zd* mes_host, const zd* otrue, DbimStats& st) {
  if (!ok()) return;
  auto t_start = std::chrono::steady_clock::now();
  k_incident<<<dim3(nblk_, T_), dim3(kBlock), 0, eng_.stream()>>>(inc_, pos_, tx_, N_, k0_, eng_.plan().cfg.radius);
  start(mes_host, otrue, st);
  st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
}

// Inversion start from measurements (mom.cpp:222-276): zero object, total
// field = incident field, mesd = -mes, first gradient and direction.
void Dbim::start(const zd* mes_host, const zd* otrue, DbimStats& st) {
  hipStream_t s = eng_.stream();
  const long N = N_;
  std::vector<zd> mes(mes_host, mes_host + (size_t)T_ * R_);
  DB_HIP(hipMemcpyAsync(mes_, mes.data(), sizeof(zd) * mes.size(), hipMemcpyHostToDevice, s));
  adj_ok_ = dtot_ok_ = false;
  objnorm_ = 0;
  if (otrue) {
    DB_HIP(hipMemcpyAsync(omes_, otrue, sizeof(zd) * N, hipMemcpyHostToDevice, s));
    for (long n = 0; n < N; n++) objnorm_ += otrue[n].x * otrue[n].x + otrue[n].y * otrue[n].y;
  } else {
    DB_HIP(hipMemsetAsync(omes_, 0, sizeof(zd) * N, s));
  }
  if (!(objnorm_ > 0)) objnorm_ = 1;
  mesnorm_ = 0;
  for (auto& v : mes) mesnorm_ += v.x * v.x + v.y * v.y;
  // initial guess: no object, total field = incident field, mesd = -mes (mom.cpp:224-229)
  for (auto& v : mes) v = mkc<double>(-v.x, -v.y);
  DB_HIP(hipMemcpyAsync(mesd_, mes.data(), sizeof(zd) * mes.size(), hipMemcpyHostToDevice, s));
  DB_HIP(hipMemsetAsync(o_, 0, sizeof(zd) * N, s));
  DB_HIP(hipMemsetAsync(g_, 0, sizeof(zd) * N, s));
  DB_HIP(hipMemsetAsync(save_, 0, sizeof(zd) * N, s));
  DB_HIP(hipMemcpyAsync(tot_, inc_, sizeof(zd) * N * T_, hipMemcpyDeviceToDevice, s));
  gradient(st);
  metrics(st);
  DB_HIP(hipMemcpyAsync(ot_, od_, sizeof(zd) * N, hipMemcpyDeviceToDevice, s));
  DB_HIP(hipMemcpyAsync(od0_, od_, sizeof(zd) * N, hipMemcpyDeviceToDevice, s));
  DB_HIP(hipStreamSynchronize(s));
  iter_ = 0;
  st.matvecs = eng_.matvecs;
}

void Dbim::restart() {
  if (!ok()) return;
  hipStream_t s = eng_.stream();
  const long N = N_;
  const size_t vb = sizeof(zd) * (size_t)N;
  k_neg<<<dim3((unsigned)std::min<long>(1024, ((long)T_ * R_ + kBlock - 1) / kBlock)), dim3(kBlock), 0, s>>>(mes_, mesd_, (long)T_ * R_);
  DB_HIP(hipMemsetAsync(o_, 0, vb, s));
  DB_HIP(hipMemsetAsync(g_, 0, vb, s));
  DB_HIP(hipMemsetAsync(save_, 0, vb, s));
  DB_HIP(hipMemcpyAsync(tot_, inc_, vb * T_, hipMemcpyDeviceToDevice, s));
  DB_HIP(hipMemcpyAsync(od_, od0_, vb, hipMemcpyDeviceToDevice, s));
  DB_HIP(hipMemcpyAsync(ot_, od0_, vb, hipMemcpyDeviceToDevice, s));
  DB_HIP(hipStreamSynchronize(s));
  adj_ok_ = dtot_ok_ = false;
  iter_ = 0;
}

void Dbim::iterate(DbimStats& st) {
  if (!ok()) return;
  auto t_start = std::chrono::steady_clock::now();
  long long mv0 = eng_.matvecs, mv32 = eng_.matvecs32;
  hipStream_t s = eng_.stream();
  const long N = N_;
  dim3 blk(kBlock);
  // Gauss-Newton step length from the Frechet derivative along ot (mom.cpp:281-316)
  for (int t0 = 0; t0 < T_; t0 += B_) {
    int nb = std::min(B_, T_ - t0);
    dim3 grid(nblk_, nb);
    zd* wy = rxw_ ? rxw_ + (size_t)t0 * N : wy_;
    k_deriv_src<<<grid, blk, 0, s>>>(tot_ + (size_t)t0 * N, ot_, wy, N, k0_);
    // opt_.products picks the operator: the FP32 one has a relative error of 1.3e-6 - 2.4e-6
    // against the FP64 one (the order of the solves' tolerance), most of it from the FP32 sums
    // of P2M and M2L, amplified by the leaf-level translations
    eng_.mlfma(wy, wc_, nb, nullptr, opt_.products);
    zd* dx = dtot_ ? dtot_ + (size_t)t0 * N : wx_;  // kept for the first-order field start
    if (opt_.solve_start == 1 && dtot_ok_) {
      solve_from(wx_, dx, wc_, false, nb, st);
    } else {
      DB_HIP(hipMemsetAsync(dx, 0, sizeof(zd) * N * nb, s));
      solve(dx, wc_, false, nb, true, st);
    }
    k_add_gx<<<grid, blk, 0, s>>>(wy, g_, dx, N);
    if (!rxw_) eng_.aggregate(wy, rhs_ + (size_t)t0 * R_, nb);
  }
  if (rxw_) eng_.aggregate(rxw_, rhs_, T_);
  dtot_ok_ = opt_.solve_start == 1;
  std::vector<zd> rhs((size_t)T_ * R_), mesd((size_t)T_ * R_);
  DB_HIP(hipMemcpyAsync(rhs.data(), rhs_, sizeof(zd) * rhs.size(), hipMemcpyDeviceToHost, s));
  DB_HIP(hipMemcpyAsync(mesd.data(), mesd_, sizeof(zd) * mesd.size(), hipMemcpyDeviceToHost, s));
  DB_HIP(hipStreamSynchronize(s));
  if (!ok()) return;
  double num = 0, den = 0;
  for (size_t m = 0; m < rhs.size(); m++) {
    num += mesd[m].x * rhs[m].x + mesd[m].y * rhs[m].y;  // Re(conj(mesd) rhs)
    den += rhs[m].x * rhs[m].x + rhs[m].y * rhs[m].y;
  }
  double alpha = num / den;
  st.alpha = alpha;
  k_axpy<<<dim3(nblk_), blk, 0, s>>>(o_, ot_, mkc<double>(-alpha, 0.0), N);
  k_weight<<<dim3(nblk_), blk, 0, s>>>(o_, g_, N, k0_);
  // total-field update for the new contrast (mom.cpp:322-336)
  DB_HIP(hipStreamSynchronize(s));
  const auto t_field = std::chrono::steady_clock::now();
  const long long it_field = st.solve_iters;
  for (int t0 = 0; t0 < T_; t0 += B_) {
    int nb = std::min(B_, T_ - t0);
    zd* tot = tot_ + (size_t)t0 * N;
    bool zero = opt_.field_start == 0;
    if (zero) DB_HIP(hipMemsetAsync(tot, 0, sizeof(zd) * N * nb, s));
    // o moved by -alpha ot, so to first order tot moves by -alpha dtot
    if (opt_.field_start >= 2)
      k_axpy_rows<<<dim3(nblk_, nb), blk, 0, s>>>(tot, dtot_ + (size_t)t0 * N, mkc<double>(-alpha, 0.0), N);
    if (opt_.field_start == 3) {
      learned_solve(tot, t0, nb, alpha, st);
    } else {
      solve(tot, inc_ + (size_t)t0 * N, false, nb, zero, st);
    }
    zd* wx = rxw_ ? rxw_ + (size_t)t0 * N : wx_;
    k_rowmul<<<dim3(nblk_, nb), blk, 0, s>>>(tot, g_, wx, N);
    if (!rxw_) eng_.aggregate(wx, rhs_ + (size_t)t0 * R_, nb);
  }
  if (rxw_) eng_.aggregate(rxw_, rhs_, T_);
  DB_HIP(hipStreamSynchronize(s));
  st.field_iters = st.solve_iters - it_field;
  st.field_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_field).count();
  DB_HIP(hipMemcpyAsync(rhs.data(), rhs_, sizeof(zd) * rhs.size(), hipMemcpyDeviceToHost, s));
  std::vector<zd> mes((size_t)T_ * R_);
  DB_HIP(hipMemcpyAsync(mes.data(), mes_, sizeof(zd) * mes.size(), hipMemcpyDeviceToHost, s));
  DB_HIP(hipStreamSynchronize(s));
  for (size_t m = 0; m < rhs.size(); m++) mesd[m] = csub(rhs[m], mes[m]);
  DB_HIP(hipMemcpyAsync(mesd_, mesd.data(), sizeof(zd) * mesd.size(), hipMemcpyHostToDevice, s));
  DB_HIP(hipMemcpyAsync(save_, od_, sizeof(zd) * N, hipMemcpyDeviceToDevice, s));
  gradient(st);
  metrics(st);
  // Polak-Ribiere (mom.cpp:383-399): beta = sum od conj(od - save) / sum |save|^2
  double s1r = 0, s1i = 0, s2 = 0;
  for (int j = 0; j < nblk_; j++) {
    s1r += part_h_[1 * nblk_ + j];
    s1i += part_h_[2 * nblk_ + j];
    s2 += part_h_[3 * nblk_ + j];
  }
  zd beta = mkc<double>(s1r / s2, s1i / s2);
  st.beta_re = beta.x;
  st.beta_im = beta.y;
  k_newdir<<<dim3(nblk_), blk, 0, s>>>(ot_, od_, beta, N);
  DB_HIP(hipStreamSynchronize(s));
  iter_++;
  st.matvecs = eng_.matvecs - mv0;
  st.matvecs32 = eng_.matvecs32 - mv32;
  st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
}

void Dbim::set_warmnet(WarmNet* net) {
  net_ = net;
  if (!net || !ok()) return;
  const Plan& P = eng_.plan();
  const long side = (long)P.side;
  if (side % WarmNet::kTile != 0 || side * side > net->max_pixels()) {
    err_ = "warm-start network workspace does not fit this domain (side multiple of 16 needed)";
    return;
  }
  if (!r2t_) {
    r2t_ = (int*)dalloc(sizeof(int) * N_);
    wdir_ = (zd*)dalloc(sizeof(zd) * (size_t)N_ * B_);
    inc_rms_ = (double*)dalloc(sizeof(double) * 2 * T_);
    rowpart_ = (double*)dalloc(sizeof(double) * nblk_ * B_);
    step_rms_ = (double*)dalloc(sizeof(double) * B_);
    out_scale_ = (double*)dalloc(sizeof(double) * B_);
    DB_HIP(hipHostMalloc((void**)&rowpart_h_, sizeof(double) * nblk_ * B_, hipHostMallocDefault));
    DB_HIP(hipHostMalloc((void**)&step_rms_h_, sizeof(double) * B_, hipHostMallocDefault));
    if (!ok()) return;
    DB_HIP(hipMemcpy(r2t_, P.raster2tree.data(), sizeof(int) * N_, hipMemcpyHostToDevice));
  }
  net_chunk_ = std::min(net->max_batch(), B_);
  rms_ok_ = false;
}

// rms of every transmitter's incident field (the network's field scale):
// inc_rms_[t] and its inverse inc_rms_[T + t]
void Dbim::incident_rms() {
  hipStream_t s = eng_.stream();
  std::vector<double> h(2 * T_);
  for (int t = 0; t < T_; t++) {
    k_sums<<<dim3(nblk_), dim3(kBlock), 0, s>>>(inc_ + (size_t)t * N_, inc_ + (size_t)t * N_, inc_ + (size_t)t * N_,
                                                inc_ + (size_t)t * N_, N_, part_);
    DB_HIP(hipMemcpyAsync(part_h_, part_, sizeof(double) * kSums * nblk_, hipMemcpyDeviceToHost, s));
    DB_HIP(hipStreamSynchronize(s));
    double e = 0;
    for (int j = 0; j < nblk_; j++) e += part_h_[4 * nblk_ + j];  // |o|^2 slot: sum |inc_t|^2
    const double r = std::sqrt(e / N_);
    h[t] = r > 0 ? r : 1.0;
    h[T_ + t] = 1.0 / h[t];
  }
  DB_HIP(hipMemcpy(inc_rms_, h.data(), sizeof(double) * 2 * T_, hipMemcpyHostToDevice));
  rms_ok_ = ok();
}

// Learned start (field_start 3): tot already holds the first-order start
// tot_prev - alpha dtot; per chunk of transmitters, pack (o, inc_t, tot_t,
// -alpha dtot_t) into raster images, run the network, scatter its correction
// back to tree order, then one blended solve for the batch.
void Dbim::learned_solve(zd* tot, int t0, int nb, double alpha, DbimStats& st) {
  if (!net_ || net_chunk_ < 1) {
    if (err_.empty()) err_ = "field_start 3 needs a warm-start network (isg_dbim_set_warmnet)";
    return;
  }
  if (!rms_ok_) incident_rms();
  hipStream_t s = eng_.stream();
  const long N = N_;
  const int side = eng_.plan().side;
  const zd* inc = inc_ + (size_t)t0 * N;
  const zd* dtot = dtot_ + (size_t)t0 * N;
  auto tn = std::chrono::steady_clock::now();
  // output scale of each transmitter: the rms of its first-order step alpha dtot_t
  // (the network predicts the correction in units of the step it refines)
  k_row_sq<<<dim3(nblk_, nb), dim3(kBlock), 0, s>>>(dtot, N, rowpart_);
  DB_HIP(hipMemcpyAsync(rowpart_h_, rowpart_, sizeof(double) * nblk_ * nb, hipMemcpyDeviceToHost, s));
  DB_HIP(hipStreamSynchronize(s));
  for (int t = 0; t < nb; t++) {
    double e = 0;
    for (int j = 0; j < nblk_; j++) e += rowpart_h_[(size_t)t * nblk_ + j];
    step_rms_h_[t] = std::fabs(alpha) * std::sqrt(e / N);
  }
  DB_HIP(hipMemcpyAsync(step_rms_, step_rms_h_, sizeof(double) * nb, hipMemcpyHostToDevice, s));
  NetScales sc;
  for (int c = 0; c < WarmNet::kInCh; c++) sc.in[c] = net_->scales()[c];
  for (int c0 = 0; c0 < nb; c0 += net_chunk_) {
    int nc = std::min(net_chunk_, nb - c0);
    float* x32 = net_->packed_input(nc, side, side);
    if (!x32) {
      if (err_.empty()) err_ = net_->error();
      return;
    }
    k_net_in<<<dim3(nblk_, nc), dim3(kBlock), 0, s>>>(o_, inc + (size_t)c0 * N, tot + (size_t)c0 * N,
                                                       dtot + (size_t)c0 * N, alpha, inc_rms_ + T_ + t0 + c0, r2t_,
                                                       side, sc, x32);
    net_->forward_packed(nc, side, side, s);
    if (!net_->ok()) {
      if (err_.empty()) err_ = net_->error();
      return;
    }
    // correction = packed output * output scale * rms of the transmitter's first-order step
    k_scale_rows<<<dim3(1), dim3(kBlock), 0, s>>>(step_rms_ + c0, (double)net_->scales()[WarmNet::kInCh], nc,
                                                   out_scale_);
    k_net_dir<<<dim3(nblk_, nc), dim3(kBlock), 0, s>>>(net_->packed_output(), out_scale_, r2t_, side,
                                                        wdir_ + (size_t)c0 * N);
  }
  DB_HIP(hipStreamSynchronize(s));
  st.net_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - tn).count();
  std::vector<SolveStat> ss;
  std::vector<zd> lam;
  std::vector<double> ratio;
  eng_.bicgs_blend(tot, wdir_, g_, 0, inc, false, nb, opt_.tol, opt_.maxit, ss, opt_.precision, &lam, false,
                   &ratio);
  double la = 0, ra = 0;
  for (auto& l : lam) la += std::sqrt(l.x * l.x + l.y * l.y);
  for (double r : ratio) ra += r;
  st.lambda_abs += lam.empty() ? 0 : la / T_;
  st.start_ratio += ratio.empty() ? 0 : ra / T_;
  for (auto& q : ss) {
    st.solves++;
    st.solve_iters += q.iters;
    if (q.status != 0) st.not_converged++;
  }
}

bool Dbim::get(int which, std::vector<zd>& out) {
  const zd* src = nullptr;
  size_t n = 0;
  switch (which) {
    case 0: src = o_; n = N_; break;
    case 1: src = od_; n = N_; break;
    case 2: src = ot_; n = N_; break;
    case 3: src = mes_; n = (size_t)T_ * R_; break;
    case 4: src = mesd_; n = (size_t)T_ * R_; break;
    case 5: src = inc_; n = (size_t)T_ * N_; break;
    case 6: src = tot_; n = (size_t)T_ * N_; break;
    case 7: src = omes_; n = N_; break;
    case 8: src = dtot_; n = (size_t)T_ * N_; break;  // null unless field_start >= 2 or solve_start 1
    default: return false;
  }
  if (!src) return false;
  out.resize(n);
  DB_HIP(hipStreamSynchronize(eng_.stream()));
  DB_HIP(hipMemcpy(out.data(), src, sizeof(zd) * n, hipMemcpyDeviceToHost));
  return ok();
}

bool Dbim::set(int which, const std::vector<zd>& in) {
  zd* dst = nullptr;
  size_t n = 0;
  switch (which) {
    case 0: dst = o_; n = N_; break;
    case 1: dst = od_; n = N_; break;
    case 2: dst = ot_; n = N_; break;
    case 4: dst = mesd_; n = (size_t)T_ * R_; break;
    case 6: dst = tot_; n = (size_t)T_ * N_; break;
    default: return false;
  }
  if (in.size() != n) return false;
  // restored state: the stored previous solutions belong to another history
  adj_ok_ = dtot_ok_ = false;
  DB_HIP(hipStreamSynchronize(eng_.stream()));
  DB_HIP(hipMemcpy(dst, in.data(), sizeof(zd) * n, hipMemcpyHostToDevice));
  if (which == 0) {
    k_weight<<<dim3(nblk_), dim3(kBlock), 0, eng_.stream()>>>(o_, g_, N_, k0_);
    DB_HIP(hipStreamSynchronize(eng_.stream()));
  }
  return ok();
}

}  // namespace isg
