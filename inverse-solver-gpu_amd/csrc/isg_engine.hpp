// isg_engine.hpp -- single-GPU MLFMA / BiCGSTAB engine (host side).
//
// One Engine owns one operator (an isg_plan), its device tables, a HIP stream
// and the device work buffers for up to `max_rhs` simultaneous right-hand
// sides.  Every vector is complex, interleaved, in tree order, laid out
// [rhs][N] in HBM; multipoles and locals of level i are [rhs][4^i][S_i].
//
// Precision: the solver vectors and scalars are always FP64.  The operator
// runs either in FP64 (plain) or, in the mixed path, with FP32 tables,
// multipoles and FP32 matrix-core arithmetic; a mixed solve is followed by
// FP64 true-residual checks and defect-correction solves until the FP64
// relative residual meets the tolerance.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "isg_device.hpp"
#include "isg_plan.hpp"

namespace isg {

enum { PREC_FP64 = 0, PREC_MIXED = 1 };

// Size of the static device scratch of the allocation-free solve (one per GPU).
constexpr size_t kStaticScratchBytes = (size_t)16 << 20;

struct EngineConfig {
  PlanConfig plan;
  int max_rhs = 1;        // right-hand sides solved together
  int precision = 0;      // PREC_MIXED also builds the FP32 operator
  double mixed_margin = 0.5;  // FP32 solve target = mixed_margin * tol before the FP64 check
  bool graph = true;          // replay each BiCGSTAB iteration as one captured HIP graph
                              // (not while live kernel timing is on: HIP events recorded
                              // by graph nodes carry no timestamps)
  int rx_rhs = 0;             // finite-radius receivers: right-hand sides one aggregate /
                              // aggregateh call may take (at least max_rhs), so that a
                              // caller solving one transmitter at a time can still apply
                              // the dense receiver operator once for all of them
};

// Result of one BiCGSTAB solve (one right-hand side).
struct SolveStat {
  int status;   // 0 converged, 1 not converged, 2 rho breakdown, 3 omega breakdown
  int iters;    // BiCGSTAB iterations, defect-correction solves included
  double relres;  // final relative residual (FP64 true residual for mixed solves)
};

// Device-resident BiCGSTAB scalars of one right-hand side.
struct BiState {
  zd rho, rho1, alpha, omega, beta;
  double bnrm, err, tol;
  int iter, status, active, first;
};


// Device copies of the operator tables and the multipole/local stores in
// precision T.
template <typename T>
struct DevOp {
  cplx<T>* near = nullptr;     // [9][64][64]
  cplx<T>* coeffT = nullptr;   // [64][S_leaf]
  cplx<T>* basisT = nullptr;   // [S_leaf][64]
  std::vector<T*> interpT, anterpT, band_m2m, band_l2l;  // band blocks [nrb][KB][16]
  std::vector<int*> bs0_m2m, bs0_l2l;                    // window start of each band block
  std::vector<cplx<T>*> shiftM, shiftL, trans;
  T* rxT = nullptr;
  T* rxaT = nullptr;
  cplx<T>* agm = nullptr;      // [level][rhs][4^i S_i]
  cplx<T>* agl = nullptr;
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
  // right-hand sides one aggregate / aggregateh call accepts
  int rx_rhs() const { return cfg_.plan.radius > 0 ? std::max(cfg_.max_rhs, cfg_.rx_rhs) : cfg_.max_rhs; }
  bool has_mixed() const { return has32_; }

  // r = G x for nrhs right-hand sides (device pointers, FP64, [nrhs][N]).
  // `act` (device, may be null) masks right-hand sides out.
  void mlfma(const zd* x, zd* r, int nrhs, const int* act = nullptr, int prec = PREC_FP64);
  // r = G x with the FP32 operator on FP32 vectors (the mixed solver's matvec)
  void mlfma_f32(const cplx<float>* x, cplx<float>* r, int nrhs);
  // far-field pattern at the receivers: r [nrhs][numrx]  (FP64)
  void aggregate(const zd* x, zd* r, int nrhs);
  // receiver back-projection: x [nrhs][N] from r [nrhs][numrx]  (FP64)
  void aggregateh(const zd* r, zd* x, int nrhs);

  // BiCGSTAB on (I - G diag(g)) x = b (herm=false) or its adjoint
  // (I - diag(conj g) G^H) (herm=true), x in/out [nrhs][N].  g is [N] when
  // g_stride == 0 (shared contrast) or [nrhs][N] when g_stride == N.
  // x_zero tells the solver that x is all zeros (skips the first matvec).
  void bicgs(zd* x, const zd* g, long g_stride, const zd* b, bool herm, int nrhs, bool x_zero, double tol,
             int maxit, std::vector<SolveStat>& out, int prec = PREC_FP64);
  // The same solve started from x + lambda d, lambda chosen per system to
  // minimise |b - A (x + lambda d)| (so the start is never worse than x); two
  // operator applications replace the usual one (one when x_zero: x holds
  // zeros and r_x = b).  lam (host, may be null) receives lambda, ratio (may
  // be null) |r_x - lambda A d| / |r_x|.
  void bicgs_blend(zd* x, const zd* d, const zd* g, long g_stride, const zd* b, bool herm, int nrhs, double tol,
                   int maxit, std::vector<SolveStat>& out, int prec, std::vector<zd>* lam, bool x_zero = false,
                   std::vector<double>* ratio = nullptr);

  // bicgs() without device allocations: the solver's partial sums, scalars,
  // flags, tolerances and blend coefficients live in this GPU's static scratch
  // array (kStaticScratchBytes, shared by every engine on the GPU, so calls
  // that overlap on one GPU race); iterations are launched kernel by kernel.
  // Needs static_scratch_bytes(nrhs) <= kStaticScratchBytes and out with
  // capacity for nrhs entries if the call must not touch the host heap either.
  void bicgs_static(zd* x, const zd* g, long g_stride, const zd* b, bool herm, int nrhs, bool x_zero, double tol,
                    int maxit, std::vector<SolveStat>& out, int prec = PREC_FP64);
  size_t static_scratch_bytes(int nrhs) const;

  // The FP64 product G prep(x) (prep(x) = g x forward, conj(x) adjoint) that
  // the last true-residual check of a mixed-precision solve computed for its
  // final iterate x, nrhs rows; null unless that check is the latest work on
  // the operator output and was made for this x, nrhs and direction.
  const zd* checked_product(const zd* x, int nrhs, bool herm) const;

  // multipoles (which=0) or locals (1) of `level` for one rhs, FP64 store
  void stage(int which, int level, int rhs, std::vector<zd>& out);

  // counters
  long long matvecs = 0;       // single-rhs operator applications
  long long matvecs32 = 0;     // ... of which with the FP32 operator
  long long solves = 0, solve_iters = 0, refinements = 0;

  void launch_prep(const zd* x, const zd* g, long g_stride, zd* y, bool herm, int nrhs);

  // Live per-kernel timing with HIP events on the engine stream.
  // KT_NEAR/KT_M2L: FP64 operator launches; KT_NEAR32/KT_M2L32: FP32 ones;
  // KT_P2M, KT_INTERP (M2M + L2L), KT_RX (receiver operator): either precision.
  enum { KT_NEAR = 0, KT_M2L = 1, KT_VEC = 2, KT_NEAR32 = 3, KT_M2L32 = 4, KT_P2M = 5, KT_INTERP = 6, KT_RX = 7,
         KT_COUNT = 8 };
  void profile(bool on);
  void profile_read(int id, double& ms, long long& launches);

  // Kernel-variant choices per operator precision and batch class (nrhs < 16
  // or >= 16); -1 = the built-in rule.  Stages: near/P2M tile width (16 or 32
  // boxes), M2L tile variant per level (kM2LVariants), M2M/L2L kernel kind per
  // level (0 band MFMA, 1 chunked LDS, 2 direct, 3 / 4 band MFMA with 16 / 32
  // columns per workgroup).
  enum { ST_NEAR = 0, ST_P2M = 1, ST_M2L = 2, ST_M2M = 3, ST_L2L = 4, ST_LEAF = 5, ST_COUNT = 6 };
  struct Tuning {
    int near_nb = -1, p2m_nb = -1;
    int leaf = -1;  // 1: leaf M2L and the last L2L as one kernel (k_m2l_leaf), 0 / -1: separate
    std::vector<int> m2l, m2m, l2l;
  };
  static int batch_class(int nrhs) { return nrhs >= 16 ? 1 : 0; }
  // Times every admissible variant of every stage with HIP events (reps
  // launches each, on zero data) and records the fastest for (each operator
  // precision, batch_class(nrhs)).
  void autotune(int nrhs, int reps);
  bool tuning_set(int prec, int cls, int stage, int level, int value);
  int tuning_get(int prec, int cls, int stage, int level) const;
  bool tuning_save(const std::string& path);
  bool tuning_load(const std::string& path);  // entries for other tree depths are skipped

 private:
  struct KTimer {
    std::vector<hipEvent_t> ev0, ev1;
    std::vector<char> starts;  // interval i opens a launch set
    int used = 0;
    double ms = 0;
    long long launches = 0;
  };
  // Launch options of one interpolation level, fixed at set-up; the dense
  // GEMM is used when it gives enough workgroups, else the chunked kernel,
  // else the direct-from-cache kernel.
  struct InterpShape {
    int ncb = 0;            // band MFMA: column blocks of 16 (0: unavailable)
    int kb = 0;             // band MFMA: window length of a 16-row block
    size_t lds_band = 0;
    bool chunk_ok = false;  // chunked LDS kernel usable
    int g = 1;              // chunked: parents per workgroup
    int kch = 256;          // chunked: target samples per workgroup
    size_t lds_chunk = 0;
  };

  template <typename T>
  DevOp<T>& ops();
  template <typename T>
  void upload(DevOp<T>& d);
  template <typename T>
  void plan_interp(DevOp<T>& d, std::vector<InterpShape>& m2m, std::vector<InterpShape>& l2l);
  // split-FP16 near-field and L2P tables of the FP32 operator (k_near_h16)
  void build_split16();
  // operator with tables / multipoles of type T on vectors of type TI
  template <typename T, typename TI>
  void apply(const cplx<TI>* y, cplx<TI>* r, int nrhs, const int* act, const NearPost<TI>* post = nullptr);
  // `force` >= 0 overrides the tuned / built-in choice (autotuner)
  template <typename T, typename TI>
  // prep (mode 3): y is the iterate x and P2M loads prep(x) (FP64 only)
  void launch_p2m(const cplx<TI>* y, int nrhs, const int* act, int force = -1, const NearPost<TI>* prep = nullptr);
  template <typename T>
  void launch_m2m(int i, int nrhs, const int* act, int force = -1);
  template <typename T>
  void launch_l2l(int i, bool accumulate, int nrhs, const int* act, int force = -1);
  template <typename T>
  void launch_m2l(int i, int nrhs, const int* act, int force = -1);
  template <typename T>
  void launch_m2l_leaf(int nrhs, const int* act);
  template <typename T>
  bool leaf_fuse_ok() const;
  template <typename T, typename TI>
  // returns the number of workgroups per right-hand side (partial sums of a post step)
  int launch_near(const cplx<TI>* y, cplx<TI>* r, int nrhs, const int* act, bool with_near, int force = -1,
                  const NearPost<TI>* post = nullptr);
  template <typename T>
  void autotune_prec(int nrhs, int reps);
  template <typename T>
  bool m2l_ok(int variant, int level) const;
  int interp_kind(const InterpShape& sh, int np, int nrhs) const;
  bool kind_ok(const InterpShape& sh, int kind) const;
  Tuning& tune(int prec, int nrhs) { return tune_[prec == PREC_MIXED ? 1 : 0][batch_class(nrhs)]; }
  std::vector<InterpShape>& m2m_shapes(bool fp32) { return fp32 ? m2m32_ : m2m64_; }

  void apply_mlfma(const zd* y, zd* r, int nrhs, const int* act, int prec);
  void device_tables(DevOp<double>& d);
  void setup_receivers();
  void rx_apply(const zd* x, zd* r, int nrhs);
  void rx_apply_h(const zd* r, zd* x, int nrhs);
  // solver work vectors of storage type V (FP32 for the mixed solver)
  template <typename V>
  cplx<V>* vb(int k);
  template <typename V>
  void apply_vec(const cplx<V>* y, cplx<V>* r, int nrhs, const int* act, int prec,
                 const NearPost<V>* post = nullptr);
  // r0_in_res: the initial residual is already in vb<V>(V_RES) (blended start)
  void solve(zd* x, const zd* g, long gs, const zd* b, bool herm, int nrhs, bool x_zero, double tol, int maxit,
             std::vector<SolveStat>& out, int prec, bool r0_in_res);
  template <typename V>
  void run_bicgs(zd* x, const zd* g, long gs, const zd* b, bool herm, int nrhs, bool x_zero, int maxit, int prec,
                 const cplx<V>* r0);
  template <typename V>
  void blend(zd* x, const zd* d, const zd* g, long gs, const zd* b, bool herm, int nrhs, int prec, bool x_zero);
  void true_residual(const zd* x, const zd* g, long gs, const zd* b, bool herm, int nrhs, std::vector<double>& rn,
                     std::vector<double>& bn);
  // new_set false: the interval adds to the previous launch set's time
  void prof_begin(int id, bool new_set = true);
  void prof_end(int id);
  void harvest();
  template <typename V>
  void bicgs_body(zd* x, const zd* g, long gs, bool herm, int nrhs, int prec);
  void* dalloc(size_t bytes);

  EngineConfig cfg_;
  Plan P_;
  std::string err_;
  hipStream_t stream_ = nullptr;
  bool has32_ = false;
  // what the operator output buffer holds after a mixed solve's last check
  const zd* checked_x_ = nullptr;
  int checked_n_ = 0;
  bool checked_h_ = false;
  DevOp<double> op64_;
  DevOp<float> op32_;
  // k_near_h16 tables ([9 x 2][4][4][64] and [ceil(S / 32)][4][4][64] groups
  // of 8 FP16) and their power-of-two scale exponents
  void* near16_ = nullptr;
  void* bas16_ = nullptr;
  int near16_et_ = 0, bas16_eb_ = 0;
  // k_p2m_h16 table ([2][4][4][16 ceil(S / 16)] groups of 8 FP16; null for S > 96) and its exponent
  void* p2m16_ = nullptr;
  int p2m16_ec_ = 0;
  std::vector<InterpShape> m2m64_, l2l64_, m2m32_, l2l32_;  // per parent level
  std::vector<int*> interp0_, anterp0_;
  int* rx0_ = nullptr;
  int* rxaI_ = nullptr;
  std::vector<size_t> lvl_off_;   // offset of level i in the stores (rhs blocks included)
  std::vector<size_t> lvl_size_;  // 4^i * S_i
  std::vector<zd*> vec_;          // solver vectors, each [max_rhs][N]
  std::vector<cplx<float>*> vec32_;  // FP32 copies of the work vectors (mixed solver)
  // finite-radius receivers: dense operator K[m][n] = (i/4) H0(k0 |rho_m - r_n|) res^2,
  // stored in chunks of rx_kc_ unknowns, rxm_[b][m][j] = K[m][b rx_kc_ + j], and applied
  // with rocBLAS: split-K partial products rxpart_, inputs copied into the same chunking
  // (rxin_[b][t][j]), conjugated receiver values rxc_
  zd* rxm_ = nullptr;
  zd* rxpart_ = nullptr;
  zd* rxc_ = nullptr;
  zd* rxin_ = nullptr;
  int rx_kc_ = 0;
  void* blas_ = nullptr;  // rocblas_handle
  double2* part_ = nullptr;       // reduction partials [max_rhs][3][part_blocks_]
  double2* part_h_ = nullptr;     // pinned host mirror, same size
  long part_blocks_ = 0;          // max(vector-kernel, near-field) workgroups per rhs and slot
  BiState* state_ = nullptr;      // [max_rhs]
  int* active_ = nullptr;         // [max_rhs]
  double* tolv_ = nullptr;        // [max_rhs] per-system tolerance
  zd* lam_ = nullptr;             // [max_rhs] blend coefficients
  double* bratio_ = nullptr;      // [max_rhs] blend residual ratios
  int* active_h_ = nullptr;       // pinned host mirrors
  BiState* state_h_ = nullptr;
  double* tol_h_ = nullptr;
  int nblk_ = 0;
  int maxit_ = 0;
  bool prof_ = false;
  KTimer kt_[KT_COUNT];
  int near_blocks_ = 0;  // workgroups per right-hand side of the last near-field launch
  Tuning tune_[2][2];  // [FP64, FP32 operator][batch class]
  std::vector<void*> allocs_;
  void* scratch_ = nullptr;          // this device's static scratch (bicgs_static)
  std::vector<double> rn_, bn_;      // true-residual norms (capacity max_rhs)
};

}  // namespace isg
