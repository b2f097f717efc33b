// isg_dbim.hpp -- distorted-Born / nonlinear-CG inversion loop on one GPU.
//
// Device-resident restatement of the reference's DBIM driver (mom.cpp:222-400,
// setup_born.cpp:8-242).  All per-transmitter fields live in HBM as
// [numtx][N] arrays; transmitters are processed in batches of up to
// Engine::max_rhs right-hand sides, so batch 1 is the reference's plain loop
// and a batch of every transmitter is one multi-RHS solve per phase.
#pragma once

#include <functional>
#include <vector>

#include "isg_engine.hpp"
#include "isg_warmnet.hpp"

namespace isg {

struct DbimOptions {
  int numtx = 16;
  double regular = 0.0;   // gradient adds regular^2 * o per transmitter (mom.cpp:267-269)
  double tol = 1e-4;      // BiCGSTAB relative residual (mom.cpp:21)
  int maxit = 300;        // BiCGSTAB iteration limit (mom.cpp:22)
  // total-field update start: 0 zero (mom.cpp:324), 1 previous total field,
  // 2 previous total field + first-order update -alpha * dtot, where dtot is the
  // Frechet derivative field the step-length phase already solved for
  // ((I - G g) dtot = G(k0^2 ot tot), mom.cpp:284-290), 3 learned: the
  // warm-start network's correction d of that first-order start, started
  // from tot_prev - alpha dtot + lambda d with the residual-minimising lambda
  int field_start = 0;
  int precision = 0;      // solver operator: PREC_FP64 or PREC_MIXED
  int products = 0;       // operator of the derivative source G (k0^2 ot tot): PREC_FP64 or PREC_MIXED
  // adjoint and derivative-field solves: 0 zero start (mom.cpp:286, :347),
  // 1 the previous DBIM iteration's solution of the same transmitter, scaled
  // by the residual-minimising lambda (never a worse start than zero)
  int solve_start = 0;
};

struct DbimStats {
  double alpha = 0;          // step length (mom.cpp:315)
  double hist[5] = {0, 0, 0, 0, 0};  // hist.txt columns (setup_born.cpp:239)
  double beta_re = 0, beta_im = 0;   // Polak-Ribiere coefficient (mom.cpp:394)
  long long solves = 0, solve_iters = 0, matvecs = 0, not_converged = 0, matvecs32 = 0;
  double seconds = 0;
  // total-field update phase (mom.cpp:322-336) alone: BiCGSTAB iterations,
  // wall time (network inference and blend included) and, for the learned
  // start, the mean |lambda| and the network time
  long long field_iters = 0;
  double field_seconds = 0, lambda_abs = 0, net_seconds = 0;
  // learned start: mean over transmitters of |r(first-order + lambda d)| / |r(first-order)|
  double start_ratio = 0;
};

class Dbim {
 public:
  Dbim(Engine& eng, const DbimOptions& opt);
  ~Dbim();
  bool ok() const { return eng_.ok() && err_.empty(); }
  const std::string& error() const { return err_.empty() ? eng_.error() : err_; }

  // Incident fields, synthetic measurements of the true contrast (tree order,
  // host array of N complex), zero initial object and the first gradient.
  void setup(const zd* otrue_host, DbimStats& st);
  // Same start from given measurements mes [numtx][numrx] (host); otrue may be
  // null (then the object-error metric is sum|o|^2).
  void setup_measured(const zd* mes_host, const zd* otrue_host, DbimStats& st);
  // One DBIM iteration (mom.cpp:277-400).
  void iterate(DbimStats& st);
  // Back to the state right after setup: zero object, total fields equal to
  // the incident fields, mesd = -mes, the first gradient and direction (kept
  // from setup, not recomputed), no previous solutions.  Device copies only.
  void restart();
  // which: 0 o, 1 od, 2 ot, 3 mes, 4 mesd, 5 inc, 6 tot, 7 true contrast; copies to host
  bool get(int which, std::vector<zd>& out);
  bool set(int which, const std::vector<zd>& in);
  int iteration() const { return iter_; }
  // Network used by field_start 3 (not owned; must outlive the iterations).
  void set_warmnet(WarmNet* net);
  // Called by setup() after each batch of synthesized measurements with
  // (transmitters done, numtx); the stream has been synchronized.
  void set_progress(std::function<void(int, int)> fn) { progress_ = std::move(fn); }

 private:
  void start(const zd* mes_host, const zd* otrue_host, DbimStats& st);
  void gradient(DbimStats& st);
  void metrics(DbimStats& st);
  void solve(zd* x, const zd* b, bool herm, int nb, bool x_zero, DbimStats& st);
  // solve into x (zeroed here) started from lambda * prev; then prev = x
  void solve_from(zd* x, zd* prev, const zd* b, bool herm, int nb, DbimStats& st);
  // tot: the first-order starts of transmitters t0..t0+nb-1 (solved in place)
  void learned_solve(zd* tot, int t0, int nb, double alpha, DbimStats& st);
  void incident_rms();
  void* dalloc(size_t bytes);

  Engine& eng_;
  DbimOptions opt_;
  std::string err_;
  int N_ = 0, R_ = 0, T_ = 0, B_ = 0;
  double k0_ = 0;
  zd *pos_ = nullptr, *tx_ = nullptr;
  zd *o_ = nullptr, *od_ = nullptr, *ot_ = nullptr, *save_ = nullptr, *omes_ = nullptr, *g_ = nullptr,
     *buff_ = nullptr;
  zd *inc_ = nullptr, *tot_ = nullptr;
  zd *mes_ = nullptr, *mesd_ = nullptr, *rhs_ = nullptr;
  zd *wy_ = nullptr, *wc_ = nullptr, *wx_ = nullptr;
  // finite-radius receivers with several transmitter batches: the receiver
  // operator's inputs (or aggregateh's outputs) of every transmitter [numtx][N],
  // so that the dense operator is applied once per phase, not once per batch
  zd* rxw_ = nullptr;
  zd* od0_ = nullptr;   // first gradient (restart)
  zd* dtot_ = nullptr;  // [numtx][N] derivative fields (field_start 2 or solve_start 1)
  zd* adj_ = nullptr;   // [numtx][N] adjoint solutions (solve_start 1)
  bool dtot_ok_ = false, adj_ok_ = false;  // the stores hold a previous iteration's solutions
  WarmNet* net_ = nullptr;
  std::function<void(int, int)> progress_;
  int net_chunk_ = 0;

  zd* wdir_ = nullptr;                             // [max_rhs][N] learned corrections
  int* r2t_ = nullptr;                             // raster -> tree index
  double* inc_rms_ = nullptr;                      // [2][numtx]: rms of inc_t, then 1 / rms
  double *rowpart_ = nullptr, *rowpart_h_ = nullptr;   // per-row partial sums [max_rhs][nblk]
  double *step_rms_ = nullptr, *step_rms_h_ = nullptr; // [max_rhs]: rms of alpha dtot_t (output scale)
  double* out_scale_ = nullptr;                        // [max_rhs]: step rms * the network's output scale
  bool rms_ok_ = false;
  double* part_ = nullptr;
  double* part_h_ = nullptr;
  int nblk_ = 0;
  double mesnorm_ = 0, objnorm_ = 0;
  int iter_ = 0;
  std::vector<void*> allocs_;
};

}  // namespace isg
