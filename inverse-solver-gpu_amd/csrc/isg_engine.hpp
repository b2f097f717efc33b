// isg_engine.hpp -- single-GPU MLFMA / BiCGSTAB / DBIM engine (host side).
//
// One Engine owns one operator (an isg_plan), its device tables, a HIP stream
// and the device work buffers for up to `max_rhs` simultaneous right-hand
// sides.  Every vector is complex, interleaved, in tree order, laid out
// [rhs][N] in HBM; multipoles and locals of level i are [rhs][4^i][S_i].
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "isg_device.hpp"
#include "isg_plan.hpp"

namespace isg {

struct EngineConfig {
  PlanConfig plan;
  int max_rhs = 1;        // right-hand sides solved together
  int precision = 0;      // 0: FP64 operator; 1: also build the FP32 operator
};

// Result of one BiCGSTAB solve (one right-hand side).
struct SolveStat {
  int status;   // 0 converged, 1 not converged, 2 rho breakdown, 3 omega breakdown
  int iters;
  double relres;
};

// Device-resident BiCGSTAB scalars of one right-hand side.
struct BiState {
  zd rho, rho1, alpha, omega, beta;
  double bnrm, err;
  int iter, status, active, first;
};

// Device copies of the operator tables in precision T.
template <typename T>
struct DevOp {
  cplx<T>* near = nullptr;     // [9][64][64]
  cplx<T>* coeffT = nullptr;   // [64][S_leaf]
  cplx<T>* basisT = nullptr;   // [S_leaf][64]
  std::vector<T*> interpT, anterpT;
  std::vector<cplx<T>*> shiftM, shiftL, trans;
  T* rxT = nullptr;
  T* rxaT = nullptr;
};

class Engine {
 public:
  explicit Engine(const EngineConfig& cfg);
  ~Engine();
  bool ok() const { return err_.empty(); }
  const std::string& error() const { return err_; }
  void set_error(const std::string& e) { if (err_.empty()) err_ = e; }
  const Plan& plan() const { return P_; }
  hipStream_t stream() const { return stream_; }
  int max_rhs() const { return cfg_.max_rhs; }

  // r = G x for nrhs right-hand sides (device pointers, FP64, [nrhs][N]).
  // `act` (device, may be null) masks right-hand sides out.
  void mlfma(const zd* x, zd* r, int nrhs, const int* act = nullptr);
  // far-field pattern at the receivers: r [nrhs][numrx]
  void aggregate(const zd* x, zd* r, int nrhs);
  // receiver back-projection: x [nrhs][N] from r [nrhs][numrx]
  void aggregateh(const zd* r, zd* x, int nrhs);

  // BiCGSTAB on (I - G diag(g)) x = b (herm=false) or its adjoint
  // (I - diag(conj g) G^H) (herm=true), x in/out [nrhs][N].  g is [N] when
  // g_stride == 0 (shared contrast) or [nrhs][N] when g_stride == N.
  // x_zero tells the solver that x is all zeros (skips the first matvec).
  void bicgs(zd* x, const zd* g, long g_stride, const zd* b, bool herm, int nrhs, bool x_zero,
             double tol, int maxit, std::vector<SolveStat>& out);

  // multipoles (which=0) or locals (1) of `level` for one rhs, FP64 host copy
  void stage(int which, int level, int rhs, std::vector<zd>& out);

  // counters
  long long matvecs = 0;       // single-rhs operator applications
  long long solves = 0, solve_iters = 0;

  // scratch vector owned by the engine ([max_rhs][N] FP64)
  zd* scratch(int k) { return vec_[k]; }

  // low-level launches used by the DBIM driver
  void launch_prep(const zd* x, const zd* g, long g_stride, zd* y, bool herm, int nrhs);

  // Live per-kernel timing with HIP events on the engine stream.
  enum { KT_NEAR = 0, KT_M2L = 1, KT_VEC = 2, KT_COUNT = 3 };
  void profile(bool on);
  void profile_read(int id, double& ms, long long& launches);

 private:
  struct KTimer {
    std::vector<hipEvent_t> ev0, ev1;
    int used = 0;
    double ms = 0;
    long long launches = 0;
  };
  void prof_begin(int id);
  void prof_end(int id);
  void harvest();
  void launch_near(const zd* y, const zd* loc, zd* r, int nrhs, const int* act, bool with_near);
  bool prof_ = false;
  KTimer kt_[KT_COUNT];

  template <typename T>
  void upload(DevOp<T>& d);
  void apply_mlfma(const zd* y, zd* r, int nrhs, const int* act);
  void reduce_finalize(int phase, int nrhs, double tol, int maxit);

  EngineConfig cfg_;
  Plan P_;
  std::string err_;
  hipStream_t stream_ = nullptr;
  DevOp<double> op64_;
  std::vector<int*> interp0_, anterp0_;
  int* rx0_ = nullptr;
  int* rxaI_ = nullptr;
  std::vector<size_t> lvl_off_;   // offset of level i in the multipole store, per rhs block
  std::vector<size_t> lvl_size_;  // 4^i * S_i
  zd* agm_ = nullptr;             // [level][rhs][4^i S_i]
  zd* agl_ = nullptr;
  std::vector<zd*> vec_;          // solver vectors, each [max_rhs][N]
  double2* part_ = nullptr;       // reduction partials [max_rhs][3][nblk]
  BiState* state_ = nullptr;      // [max_rhs]
  int* active_ = nullptr;         // [max_rhs]
  int* active_h_ = nullptr;       // pinned host mirror
  BiState* state_h_ = nullptr;    // pinned host mirror
  int nblk_ = 0;
  std::vector<void*> allocs_;
  void* dalloc(size_t bytes);
};

}  // namespace isg
