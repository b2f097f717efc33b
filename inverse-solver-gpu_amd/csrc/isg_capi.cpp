// isg_capi.cpp -- extern "C" entry points of libisg.so (declared in include/isg.h).
#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/isg.h"
#include "isg_dbim.hpp"
#include "isg_engine.hpp"
#include "isg_warmnet.hpp"

struct isg_warmnet {
  std::unique_ptr<isg::WarmNet> net;
  std::string err;
  // device staging of the host-array entry points
  float* din = nullptr;
  float* dout = nullptr;
  size_t cap = 0;  // pixels
  ~isg_warmnet() {
    if (din) (void)hipFree(din);
    if (dout) (void)hipFree(dout);
  }
};

struct isg_context {
  isg_config cfg;
  std::unique_ptr<isg::Engine> eng;
  std::unique_ptr<isg::Dbim> dbim;
  std::string err;
  // host-pointer staging buffers (device), single right-hand side
  isg::zd* dx = nullptr;
  isg::zd* dr = nullptr;
  isg::zd* dg = nullptr;
  isg::zd* drx = nullptr;
  // per-system results of isg_bicgs_dev_static (capacity max_rhs, reserved at creation)
  std::vector<isg::SolveStat> sst;
  // isg_set_progress
  isg_progress_fn progress = nullptr;
  void* progress_user = nullptr;
};

namespace {

int fail(isg_context* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

bool check(isg_context* c) {
  if (!c || !c->eng) return false;
  if (!c->eng->ok()) {
    c->err = c->eng->error();
    return false;
  }
  return true;
}

bool hip_ok(isg_context* c, hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  c->err = std::string(what) + ": " + hipGetErrorString(e);
  return false;
}

int sync(isg_context* c) {
  if (!hip_ok(c, hipStreamSynchronize(c->eng->stream()), "stream sync")) return ISG_E_DEVICE;
  return check(c) ? ISG_OK : ISG_E_DEVICE;
}

size_t vbytes(isg_context* c) { return sizeof(isg::zd) * (size_t)c->eng->plan().N; }

}  // namespace

extern "C" {

void isg_config_default(isg_config* c) {
  if (!c) return;
  c->level = 7;
  c->box = 8;
  c->res = 0.1;
  c->scale = 1.0;
  c->beta = 5;
  c->ninter = 30;
  c->numrx = 16;
  c->numtx = 16;
  c->toli = 1e-4;
  c->iti = 300;
  c->regular = 0.0;
  c->max_rhs = 1;
  c->precision = 0;
  c->mixed_margin = 0.5;
  c->graph = 1;
  c->radius = 0.0;
  c->products = 0;
  c->solve_start = 0;
  c->setup = 1;
}

int isg_create(const isg_config* cfg, isg_context** out) {
  if (!cfg || !out) return ISG_E_ARG;
  *out = nullptr;
  if (cfg->max_rhs < 1) return ISG_E_ARG;
  if (cfg->precision != 0 && cfg->precision != 1) return ISG_E_ARG;
  if (!(cfg->mixed_margin > 0 && cfg->mixed_margin <= 1)) return ISG_E_ARG;
  if (!(cfg->radius >= 0)) return ISG_E_ARG;
  if (cfg->products != 0 && !(cfg->products == 1 && cfg->precision == 1)) return ISG_E_ARG;
  if (cfg->solve_start != 0 && cfg->solve_start != 1) return ISG_E_ARG;
  if (cfg->setup != 0 && cfg->setup != 1) return ISG_E_ARG;
  isg_context* c = new isg_context();
  c->cfg = *cfg;
  isg::EngineConfig ec;
  ec.plan.level = cfg->level;
  ec.plan.box = cfg->box;
  ec.plan.res = cfg->res;
  ec.plan.scale = cfg->scale;
  ec.plan.beta = cfg->beta;
  ec.plan.ninter = cfg->ninter;
  ec.plan.numrx = cfg->numrx;
  ec.plan.radius = cfg->radius;
  ec.plan.host_tables = cfg->setup == 0;
  ec.max_rhs = cfg->max_rhs;
  ec.precision = cfg->precision;
  ec.mixed_margin = cfg->mixed_margin;
  ec.graph = cfg->graph != 0;
  // the DBIM driver applies the receiver operator once per phase for every transmitter
  ec.rx_rhs = std::max(cfg->max_rhs, cfg->numtx);
  c->eng.reset(new isg::Engine(ec));
  c->sst.reserve(cfg->max_rhs);
  if (!check(c)) {
    *out = c;  // keep the context so the caller can read the error
    return ISG_E_DEVICE;
  }
  size_t vb = vbytes(c);
  if (!hip_ok(c, hipMalloc((void**)&c->dx, vb), "hipMalloc") || !hip_ok(c, hipMalloc((void**)&c->dr, vb), "hipMalloc") ||
      !hip_ok(c, hipMalloc((void**)&c->dg, vb), "hipMalloc") ||
      !hip_ok(c, hipMalloc((void**)&c->drx, sizeof(isg::zd) * (size_t)cfg->numrx), "hipMalloc")) {
    *out = c;
    return ISG_E_DEVICE;
  }
  *out = c;
  return ISG_OK;
}

void isg_destroy(isg_context* c) {
  if (!c) return;
  c->dbim.reset();
  c->eng.reset();
  if (c->dx) (void)hipFree(c->dx);
  if (c->dr) (void)hipFree(c->dr);
  if (c->dg) (void)hipFree(c->dg);
  if (c->drx) (void)hipFree(c->drx);
  delete c;
}

const char* isg_last_error(const isg_context* c) { return c ? c->err.c_str() : "null context"; }

int isg_numunk(const isg_context* c) { return (c && c->eng) ? c->eng->plan().N : ISG_E_ARG; }
int isg_level(const isg_context* c) { return (c && c->eng) ? c->eng->plan().L : ISG_E_ARG; }
int isg_numsamp(const isg_context* c, int level) {
  if (!c || !c->eng || level < 0 || level >= c->eng->plan().L) return ISG_E_ARG;
  return c->eng->plan().nsamp[level];
}
double isg_k0(const isg_context* c) { return (c && c->eng) ? c->eng->plan().k0 : 0.0; }

int isg_geometry(const isg_context* c, double* pos, int* r2t) {
  if (!c || !c->eng) return ISG_E_ARG;
  const isg::Plan& P = c->eng->plan();
  if (pos) std::memcpy(pos, P.pos.data(), sizeof(double) * 2 * P.N);
  if (r2t) std::memcpy(r2t, P.raster2tree.data(), sizeof(int) * P.N);
  return ISG_OK;
}

void* isg_stream(const isg_context* c) { return (c && c->eng) ? (void*)c->eng->stream() : nullptr; }

int isg_mlfma(isg_context* c, const double* x, double* r) {
  if (!check(c)) return ISG_E_STATE;
  if (!x || !r) return fail(c, ISG_E_ARG, "null pointer");
  hipStream_t s = c->eng->stream();
  size_t vb = vbytes(c);
  if (!hip_ok(c, hipMemcpyAsync(c->dx, x, vb, hipMemcpyHostToDevice, s), "H2D")) return ISG_E_DEVICE;
  c->eng->mlfma(c->dx, c->dr, 1);
  if (!hip_ok(c, hipMemcpyAsync(r, c->dr, vb, hipMemcpyDeviceToHost, s), "D2H")) return ISG_E_DEVICE;
  return sync(c);
}

int isg_mlfma_many(isg_context* c, const double* x, double* r, int nrhs) {
  if (!check(c)) return ISG_E_STATE;
  if (!x || !r || nrhs < 1 || nrhs > c->cfg.max_rhs) return fail(c, ISG_E_ARG, "bad argument");
  hipStream_t s = c->eng->stream();
  size_t bytes = vbytes(c) * (size_t)nrhs;
  isg::zd *dx = nullptr, *dr = nullptr;
  if (!hip_ok(c, hipMalloc((void**)&dx, bytes), "hipMalloc") || !hip_ok(c, hipMalloc((void**)&dr, bytes), "hipMalloc")) {
    if (dx) (void)hipFree(dx);
    return ISG_E_DEVICE;
  }
  int rc = ISG_OK;
  if (!hip_ok(c, hipMemcpyAsync(dx, x, bytes, hipMemcpyHostToDevice, s), "H2D")) rc = ISG_E_DEVICE;
  if (rc == ISG_OK) {
    c->eng->mlfma(dx, dr, nrhs, nullptr, c->cfg.precision);
    if (!hip_ok(c, hipMemcpyAsync(r, dr, bytes, hipMemcpyDeviceToHost, s), "D2H")) rc = ISG_E_DEVICE;
  }
  if (rc == ISG_OK) rc = sync(c);
  (void)hipFree(dx);
  (void)hipFree(dr);
  return rc;
}

int isg_mlfmah(isg_context* c, const double* x, double* r) {
  if (!check(c)) return ISG_E_STATE;
  if (!x || !r) return fail(c, ISG_E_ARG, "null pointer");
  const int N = c->eng->plan().N;
  std::vector<double> xc(x, x + 2 * (size_t)N);
  for (int n = 0; n < N; n++) xc[2 * n + 1] = -xc[2 * n + 1];
  int rc = isg_mlfma(c, xc.data(), r);
  if (rc != ISG_OK) return rc;
  for (int n = 0; n < N; n++) r[2 * n + 1] = -r[2 * n + 1];
  return ISG_OK;
}

int isg_aggregate(isg_context* c, const double* x, double* r) {
  if (!check(c)) return ISG_E_STATE;
  if (!x || !r) return fail(c, ISG_E_ARG, "null pointer");
  hipStream_t s = c->eng->stream();
  if (!hip_ok(c, hipMemcpyAsync(c->dx, x, vbytes(c), hipMemcpyHostToDevice, s), "H2D")) return ISG_E_DEVICE;
  c->eng->aggregate(c->dx, c->drx, 1);
  if (!hip_ok(c, hipMemcpyAsync(r, c->drx, sizeof(isg::zd) * c->cfg.numrx, hipMemcpyDeviceToHost, s), "D2H"))
    return ISG_E_DEVICE;
  return sync(c);
}

int isg_aggregateh(isg_context* c, double* x, const double* r) {
  if (!check(c)) return ISG_E_STATE;
  if (!x || !r) return fail(c, ISG_E_ARG, "null pointer");
  hipStream_t s = c->eng->stream();
  if (!hip_ok(c, hipMemcpyAsync(c->drx, r, sizeof(isg::zd) * c->cfg.numrx, hipMemcpyHostToDevice, s), "H2D"))
    return ISG_E_DEVICE;
  c->eng->aggregateh(c->drx, c->dx, 1);
  if (!hip_ok(c, hipMemcpyAsync(x, c->dx, vbytes(c), hipMemcpyDeviceToHost, s), "D2H")) return ISG_E_DEVICE;
  return sync(c);
}

int isg_bicgs(isg_context* c, double* x, const double* o, const double* b, int herm, isg_solve_info* info) {
  if (!check(c)) return ISG_E_STATE;
  if (!x || !o || !b) return fail(c, ISG_E_ARG, "null pointer");
  hipStream_t s = c->eng->stream();
  size_t vb = vbytes(c);
  if (!hip_ok(c, hipMemcpyAsync(c->dx, x, vb, hipMemcpyHostToDevice, s), "H2D") ||
      !hip_ok(c, hipMemcpyAsync(c->dg, o, vb, hipMemcpyHostToDevice, s), "H2D") ||
      !hip_ok(c, hipMemcpyAsync(c->dr, b, vb, hipMemcpyHostToDevice, s), "H2D"))
    return ISG_E_DEVICE;
  std::vector<isg::SolveStat> st;
  c->eng->bicgs(c->dx, c->dg, 0, c->dr, herm != 0, 1, false, c->cfg.toli, c->cfg.iti, st, c->cfg.precision);
  if (!check(c)) return ISG_E_DEVICE;
  if (!hip_ok(c, hipMemcpyAsync(x, c->dx, vb, hipMemcpyDeviceToHost, s), "D2H")) return ISG_E_DEVICE;
  int rc = sync(c);
  if (rc != ISG_OK) return rc;
  if (info) {
    info->status = st[0].status;
    info->iters = st[0].iters;
    info->relres = st[0].relres;
  }
  return st[0].status;
}

int isg_mlfma_dev(isg_context* c, const void* x, void* r, int nrhs) {
  if (!check(c)) return ISG_E_STATE;
  if (!x || !r || nrhs < 1 || nrhs > c->cfg.max_rhs) return fail(c, ISG_E_ARG, "bad argument");
  c->eng->mlfma((const isg::zd*)x, (isg::zd*)r, nrhs);
  return check(c) ? ISG_OK : ISG_E_DEVICE;
}

int isg_synchronize(isg_context* c) {
  if (!check(c)) return ISG_E_STATE;
  return sync(c);
}

int isg_mlfma_f32_dev(isg_context* c, const void* x, void* r, int nrhs) {
  if (!check(c)) return ISG_E_STATE;
  if (!x || !r || nrhs < 1 || nrhs > c->cfg.max_rhs) return fail(c, ISG_E_ARG, "bad argument");
  c->eng->mlfma_f32((const isg::cplx<float>*)x, (isg::cplx<float>*)r, nrhs);
  return check(c) ? ISG_OK : ISG_E_DEVICE;
}

int isg_aggregate_dev(isg_context* c, const void* x, void* r, int nrhs) {
  if (!check(c)) return ISG_E_STATE;
  if (!x || !r || nrhs < 1 || nrhs > c->cfg.max_rhs) return fail(c, ISG_E_ARG, "bad argument");
  c->eng->aggregate((const isg::zd*)x, (isg::zd*)r, nrhs);
  return check(c) ? ISG_OK : ISG_E_DEVICE;
}

int isg_aggregateh_dev(isg_context* c, void* x, const void* r, int nrhs) {
  if (!check(c)) return ISG_E_STATE;
  if (!x || !r || nrhs < 1 || nrhs > c->cfg.max_rhs) return fail(c, ISG_E_ARG, "bad argument");
  c->eng->aggregateh((const isg::zd*)r, (isg::zd*)x, nrhs);
  return check(c) ? ISG_OK : ISG_E_DEVICE;
}

int isg_bicgs_dev(isg_context* c, void* x, const void* o, long o_stride, const void* b, int herm, int nrhs,
                  int x_zero, isg_solve_info* info) {
  if (!check(c)) return ISG_E_STATE;
  if (!x || !o || !b || nrhs < 1 || nrhs > c->cfg.max_rhs) return fail(c, ISG_E_ARG, "bad argument");
  if (o_stride != 0 && o_stride != c->eng->plan().N) return fail(c, ISG_E_ARG, "o_stride must be 0 or numunk");
  std::vector<isg::SolveStat> st;
  c->eng->bicgs((isg::zd*)x, (const isg::zd*)o, o_stride, (const isg::zd*)b, herm != 0, nrhs, x_zero != 0,
                c->cfg.toli, c->cfg.iti, st, c->cfg.precision);
  if (!check(c)) return ISG_E_DEVICE;
  int worst = 0;
  for (int i = 0; i < nrhs; i++) {
    if (info) {
      info[i].status = st[i].status;
      info[i].iters = st[i].iters;
      info[i].relres = st[i].relres;
    }
    if (st[i].status > worst) worst = st[i].status;
  }
  return worst;
}

int isg_bicgs_dev_static(isg_context* c, void* x, const void* o, long o_stride, const void* b, int herm, int nrhs,
                         int x_zero, isg_solve_info* info) {
  if (!check(c)) return ISG_E_STATE;
  if (!x || !o || !b || nrhs < 1 || nrhs > c->cfg.max_rhs) return fail(c, ISG_E_ARG, "bad argument");
  if (o_stride != 0 && o_stride != c->eng->plan().N) return fail(c, ISG_E_ARG, "o_stride must be 0 or numunk");
  if (c->eng->static_scratch_bytes(nrhs) > isg::kStaticScratchBytes)
    return fail(c, ISG_E_ARG, "isg_bicgs_dev_static: this problem needs " +
                                  std::to_string(c->eng->static_scratch_bytes(nrhs)) +
                                  " bytes of solver scratch, more than the static " +
                                  std::to_string(isg::kStaticScratchBytes) + "; use isg_bicgs_dev");
  c->eng->bicgs_static((isg::zd*)x, (const isg::zd*)o, o_stride, (const isg::zd*)b, herm != 0, nrhs, x_zero != 0,
                       c->cfg.toli, c->cfg.iti, c->sst, c->cfg.precision);
  if (!check(c)) return ISG_E_DEVICE;
  int worst = 0;
  for (int i = 0; i < nrhs; i++) {
    if (info) {
      info[i].status = c->sst[i].status;
      info[i].iters = c->sst[i].iters;
      info[i].relres = c->sst[i].relres;
    }
    if (c->sst[i].status > worst) worst = c->sst[i].status;
  }
  return worst;
}

int isg_stage(isg_context* c, int which, int level, int rhs, double* out) {
  if (!check(c)) return ISG_E_STATE;
  if (!out || level < 0 || level >= c->eng->plan().L) return fail(c, ISG_E_ARG, "bad argument");
  if (which != 0 && which != 1) return fail(c, ISG_E_ARG, "which must be 0 (multipoles) or 1 (locals)");
  if (rhs < 0 || rhs >= c->eng->max_rhs()) return fail(c, ISG_E_ARG, "rhs out of range [0, max_rhs)");
  std::vector<isg::zd> v;
  c->eng->stage(which, level, rhs, v);
  if (!check(c)) return ISG_E_DEVICE;
  if (v.empty()) return fail(c, ISG_E_STATE, "stage store unavailable");
  std::memcpy(out, v.data(), v.size() * sizeof(isg::zd));
  return ISG_OK;
}

int isg_counters(const isg_context* c, long long* out) {
  if (!c || !c->eng || !out) return ISG_E_ARG;
  out[0] = c->eng->matvecs;
  out[1] = c->eng->solves;
  out[2] = c->eng->solve_iters;
  out[3] = c->eng->matvecs32;
  out[4] = c->eng->refinements;
  return ISG_OK;
}

static void fill_stats(const isg::DbimStats& s, isg_dbim_stats* o) {
  if (!o) return;
  o->alpha = s.alpha;
  o->misfit = s.hist[0];
  o->object_error = s.hist[1];
  o->object_cost = s.hist[2];
  o->gradient_norm2 = s.hist[3];
  o->object_norm2 = s.hist[4];
  o->beta_re = s.beta_re;
  o->beta_im = s.beta_im;
  o->solves = s.solves;
  o->solve_iters = s.solve_iters;
  o->matvecs = s.matvecs;
  o->matvecs_fp32 = s.matvecs32;
  o->not_converged = s.not_converged;
  o->seconds = s.seconds;
  o->field_iters = s.field_iters;
  o->field_seconds = s.field_seconds;
  o->lambda_abs = s.lambda_abs;
  o->start_ratio = s.start_ratio;
  o->net_seconds = s.net_seconds;
}

int isg_dbim_setup(isg_context* c, const double* o_true, int field_start, isg_dbim_stats* st) {
  if (!check(c)) return ISG_E_STATE;
  if (!o_true || field_start < 0 || field_start > 3) return fail(c, ISG_E_ARG, "bad argument");
  isg::DbimOptions opt;
  opt.numtx = c->cfg.numtx;
  opt.regular = c->cfg.regular;
  opt.tol = c->cfg.toli;
  opt.maxit = c->cfg.iti;
  opt.field_start = field_start;
  opt.precision = c->cfg.precision;
  opt.products = c->cfg.products;
  opt.solve_start = c->cfg.solve_start;
  c->dbim.reset();
  c->dbim.reset(new isg::Dbim(*c->eng, opt));
  if (!c->dbim->ok()) return fail(c, ISG_E_DEVICE, c->dbim->error());
  isg::DbimStats s;
  if (c->progress) {
    isg_progress_fn fn = c->progress;
    void* user = c->progress_user;
    c->dbim->set_progress([fn, user](int done, int total) { fn(user, done, total); });
  }
  c->dbim->setup((const isg::zd*)o_true, s);
  if (!c->dbim->ok()) return fail(c, ISG_E_DEVICE, c->dbim->error());
  fill_stats(s, st);
  return ISG_OK;
}

int isg_dbim_setup_measured(isg_context* c, const double* mes, const double* o_true, int field_start,
                            isg_dbim_stats* st) {
  if (!check(c)) return ISG_E_STATE;
  if (!mes || field_start < 0 || field_start > 3) return fail(c, ISG_E_ARG, "bad argument");
  isg::DbimOptions opt;
  opt.numtx = c->cfg.numtx;
  opt.regular = c->cfg.regular;
  opt.tol = c->cfg.toli;
  opt.maxit = c->cfg.iti;
  opt.field_start = field_start;
  opt.precision = c->cfg.precision;
  opt.products = c->cfg.products;
  opt.solve_start = c->cfg.solve_start;
  c->dbim.reset();
  c->dbim.reset(new isg::Dbim(*c->eng, opt));
  if (!c->dbim->ok()) return fail(c, ISG_E_DEVICE, c->dbim->error());
  isg::DbimStats s;
  c->dbim->setup_measured((const isg::zd*)mes, (const isg::zd*)o_true, s);
  if (!c->dbim->ok()) return fail(c, ISG_E_DEVICE, c->dbim->error());
  fill_stats(s, st);
  return ISG_OK;
}

int isg_autotune(isg_context* c, int nrhs, int reps) {
  if (!check(c)) return ISG_E_STATE;
  if (nrhs < 1 || nrhs > c->cfg.max_rhs || reps < 1) return fail(c, ISG_E_ARG, "bad argument");
  c->eng->autotune(nrhs, reps);
  return check(c) ? ISG_OK : ISG_E_DEVICE;
}

int isg_tuning_set(isg_context* c, int precision, int batch_class, int stage, int level, int value) {
  if (!check(c)) return ISG_E_STATE;
  return c->eng->tuning_set(precision, batch_class, stage, level, value) ? ISG_OK
                                                                       : fail(c, ISG_E_ARG, "bad tuning entry");
}

int isg_tuning_get(isg_context* c, int precision, int batch_class, int stage, int level, int* value) {
  if (!check(c)) return ISG_E_STATE;
  if (!value) return fail(c, ISG_E_ARG, "null pointer");
  *value = c->eng->tuning_get(precision, batch_class, stage, level);
  return ISG_OK;
}

int isg_tuning_save(isg_context* c, const char* path) {
  if (!check(c)) return ISG_E_STATE;
  if (!path || !c->eng->tuning_save(path)) return fail(c, ISG_E_ARG, "cannot write tuning file");
  return ISG_OK;
}

int isg_tuning_load(isg_context* c, const char* path) {
  if (!check(c)) return ISG_E_STATE;
  if (!path || !c->eng->tuning_load(path)) return fail(c, ISG_E_ARG, "cannot read tuning file");
  return ISG_OK;
}

int isg_profile(isg_context* c, int enable) {
  if (!check(c)) return ISG_E_STATE;
  c->eng->profile(enable != 0);
  return check(c) ? ISG_OK : ISG_E_DEVICE;
}

int isg_profile_read(isg_context* c, int kernel, double* total_ms, long long* launches) {
  if (!check(c)) return ISG_E_STATE;
  if (kernel < 0 || kernel >= isg::Engine::KT_COUNT || !total_ms || !launches) return fail(c, ISG_E_ARG, "bad argument");
  c->eng->profile_read(kernel, *total_ms, *launches);
  return check(c) ? ISG_OK : ISG_E_DEVICE;
}

int isg_dbim_iterate(isg_context* c, isg_dbim_stats* st) {
  if (!check(c)) return ISG_E_STATE;
  if (!c->dbim) return fail(c, ISG_E_STATE, "isg_dbim_setup has not been called");
  isg::DbimStats s;
  c->dbim->iterate(s);
  if (!c->dbim->ok()) return fail(c, ISG_E_DEVICE, c->dbim->error());
  fill_stats(s, st);
  return ISG_OK;
}

int isg_set_progress(isg_context* c, isg_progress_fn fn, void* user) {
  if (!check(c)) return ISG_E_STATE;
  c->progress = fn;
  c->progress_user = user;
  return ISG_OK;
}

int isg_dbim_restart(isg_context* c) {
  if (!check(c)) return ISG_E_STATE;
  if (!c->dbim) return fail(c, ISG_E_STATE, "isg_dbim_setup has not been called");
  c->dbim->restart();
  if (!c->dbim->ok()) return fail(c, ISG_E_DEVICE, c->dbim->error());
  return ISG_OK;
}

int isg_dbim_set_warmnet(isg_context* c, isg_warmnet* n) {
  if (!check(c)) return ISG_E_STATE;
  if (!c->dbim) return fail(c, ISG_E_STATE, "isg_dbim_setup has not been called");
  if (!n || !n->net || !n->net->ok()) return fail(c, ISG_E_ARG, "no usable warm-start network");
  c->dbim->set_warmnet(n->net.get());
  if (!c->dbim->ok()) return fail(c, ISG_E_ARG, c->dbim->error());
  return ISG_OK;
}

// ---------------------------------------------------------------- warm-start network
static int net_fail(isg_warmnet* n, int code, const std::string& msg) {
  if (n) n->err = msg;
  return code;
}

static int net_check(isg_warmnet* n) {
  if (!n || !n->net) return ISG_E_STATE;
  if (!n->net->ok()) return net_fail(n, ISG_E_DEVICE, n->net->error());
  return ISG_OK;
}

int isg_warmnet_create(int max_batch, long max_pixels, unsigned long long seed, isg_warmnet** out) {
  if (!out) return ISG_E_ARG;
  *out = nullptr;
  isg_warmnet* n = new isg_warmnet();
  n->net.reset(new isg::WarmNet(max_batch, max_pixels, seed));
  *out = n;
  if (!n->net->ok()) {
    n->err = n->net->error();
    return n->err.find("no HIP device") != std::string::npos ? ISG_E_DEVICE : ISG_E_ARG;
  }
  return ISG_OK;
}

void isg_warmnet_destroy(isg_warmnet* n) { delete n; }

const char* isg_warmnet_last_error(const isg_warmnet* n) { return n ? n->err.c_str() : "null network"; }

long isg_warmnet_num_params(void) { return (long)isg::WarmNet::num_params(); }

int isg_warmnet_get_params(isg_warmnet* n, float* out) {
  if (net_check(n) != ISG_OK) return ISG_E_STATE;
  if (!out) return net_fail(n, ISG_E_ARG, "null pointer");
  std::vector<float> w;
  n->net->get_params(w);
  std::memcpy(out, w.data(), w.size() * sizeof(float));
  return net_check(n);
}

int isg_warmnet_set_params(isg_warmnet* n, const float* in) {
  if (net_check(n) != ISG_OK) return ISG_E_STATE;
  if (!in) return net_fail(n, ISG_E_ARG, "null pointer");
  std::vector<float> w(in, in + isg::WarmNet::num_params());
  n->net->set_params(w);
  return net_check(n);
}

int isg_warmnet_get_grads(isg_warmnet* n, float* out) {
  if (net_check(n) != ISG_OK) return ISG_E_STATE;
  if (!out) return net_fail(n, ISG_E_ARG, "null pointer");
  std::vector<float> g;
  n->net->get_grads(g);
  std::memcpy(out, g.data(), g.size() * sizeof(float));
  return net_check(n);
}

int isg_warmnet_set_scales(isg_warmnet* n, const float* s) {
  if (net_check(n) != ISG_OK) return ISG_E_STATE;
  if (!s) return net_fail(n, ISG_E_ARG, "null pointer");
  n->net->set_scales(s);
  return ISG_OK;
}

int isg_warmnet_get_scales(isg_warmnet* n, float* s) {
  if (net_check(n) != ISG_OK) return ISG_E_STATE;
  if (!s) return net_fail(n, ISG_E_ARG, "null pointer");
  std::memcpy(s, n->net->scales(), sizeof(float) * isg::WarmNet::kScales);
  return ISG_OK;
}

int isg_warmnet_forward_dev(isg_warmnet* n, const void* in, void* out, int batch, int H, int W) {
  if (net_check(n) != ISG_OK) return ISG_E_STATE;
  if (!in || !out) return net_fail(n, ISG_E_ARG, "null pointer");
  n->net->forward((const float*)in, (float*)out, batch, H, W, nullptr);
  if (net_check(n) != ISG_OK) return ISG_E_ARG;
  if (hipStreamSynchronize(nullptr) != hipSuccess) return net_fail(n, ISG_E_DEVICE, "stream sync");
  return net_check(n);
}

int isg_warmnet_loss_grad_dev(isg_warmnet* n, const void* in, const void* target, int batch, int H, int W,
                              double* loss) {
  if (net_check(n) != ISG_OK) return ISG_E_STATE;
  if (!in || !target || !loss) return net_fail(n, ISG_E_ARG, "null pointer");
  *loss = n->net->loss_grad((const float*)in, (const float*)target, batch, H, W, nullptr);
  return net_check(n) == ISG_OK ? ISG_OK : ISG_E_ARG;
}

static bool net_stage(isg_warmnet* n, size_t pix) {
  if (pix <= n->cap) return true;
  if (n->din) (void)hipFree(n->din);
  if (n->dout) (void)hipFree(n->dout);
  n->din = n->dout = nullptr;
  n->cap = 0;
  if (hipMalloc((void**)&n->din, pix * isg::WarmNet::kInCh * sizeof(float)) != hipSuccess ||
      hipMalloc((void**)&n->dout, pix * 2 * sizeof(float)) != hipSuccess)
    return false;
  n->cap = pix;
  return true;
}

int isg_warmnet_forward(isg_warmnet* n, const float* in, float* out, int batch, int H, int W) {
  if (net_check(n) != ISG_OK) return ISG_E_STATE;
  if (!in || !out || batch < 1 || H < 1 || W < 1) return net_fail(n, ISG_E_ARG, "bad argument");
  const size_t pix = (size_t)batch * H * W;
  if (!net_stage(n, pix)) return net_fail(n, ISG_E_DEVICE, "hipMalloc failed");
  if (hipMemcpy(n->din, in, pix * isg::WarmNet::kInCh * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    return net_fail(n, ISG_E_DEVICE, "copy to device failed");
  int rc = isg_warmnet_forward_dev(n, n->din, n->dout, batch, H, W);
  if (rc != ISG_OK) return rc;
  if (hipMemcpy(out, n->dout, pix * 2 * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    return net_fail(n, ISG_E_DEVICE, "copy to host failed");
  return ISG_OK;
}

int isg_warmnet_loss_grad(isg_warmnet* n, const float* in, const float* target, int batch, int H, int W,
                          double* loss) {
  if (net_check(n) != ISG_OK) return ISG_E_STATE;
  if (!in || !target || !loss || batch < 1 || H < 1 || W < 1) return net_fail(n, ISG_E_ARG, "bad argument");
  const size_t pix = (size_t)batch * H * W;
  if (!net_stage(n, pix)) return net_fail(n, ISG_E_DEVICE, "hipMalloc failed");
  if (hipMemcpy(n->din, in, pix * isg::WarmNet::kInCh * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(n->dout, target, pix * 2 * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    return net_fail(n, ISG_E_DEVICE, "copy to device failed");
  return isg_warmnet_loss_grad_dev(n, n->din, n->dout, batch, H, W, loss);
}

int isg_warmnet_adam(isg_warmnet* n, float lr) {
  if (net_check(n) != ISG_OK) return ISG_E_STATE;
  n->net->adam(lr, nullptr);
  if (hipStreamSynchronize(nullptr) != hipSuccess) return net_fail(n, ISG_E_DEVICE, "stream sync");
  return net_check(n);
}

int isg_warmnet_save(isg_warmnet* n, const char* path) {
  if (net_check(n) != ISG_OK) return ISG_E_STATE;
  if (!path) return net_fail(n, ISG_E_ARG, "null pointer");
  return n->net->save(path) ? ISG_OK : net_fail(n, ISG_E_ARG, n->net->error());
}

int isg_warmnet_load(isg_warmnet* n, const char* path) {
  if (net_check(n) != ISG_OK) return ISG_E_STATE;
  if (!path) return net_fail(n, ISG_E_ARG, "null pointer");
  return n->net->load(path) ? ISG_OK : net_fail(n, ISG_E_ARG, n->net->error());
}

int isg_dbim_get(isg_context* c, int which, double* out) {
  if (!check(c)) return ISG_E_STATE;
  if (!c->dbim) return fail(c, ISG_E_STATE, "isg_dbim_setup has not been called");
  if (!out) return fail(c, ISG_E_ARG, "null pointer");
  std::vector<isg::zd> v;
  if (!c->dbim->get(which, v)) return fail(c, ISG_E_ARG, "bad field or device error");
  std::memcpy(out, v.data(), v.size() * sizeof(isg::zd));
  return ISG_OK;
}

// cg_<k>.bin / cg_ref.bin (setup_born.cpp:115-121, 209-218): the contrast in
// raster order, n-th entry = tree entry unkmap[n]
int isg_dbim_write_cg(isg_context* c, const char* path, int which) {
  if (!check(c)) return ISG_E_STATE;
  if (!c->dbim) return fail(c, ISG_E_STATE, "isg_dbim_setup has not been called");
  if (!path || (which != 0 && which != 7)) return fail(c, ISG_E_ARG, "bad argument");
  std::vector<isg::zd> v;
  if (!c->dbim->get(which, v)) return fail(c, ISG_E_DEVICE, c->dbim->error());
  const std::vector<int>& r2t = c->eng->plan().raster2tree;
  std::vector<isg::zd> ras(v.size());
  for (size_t q = 0; q < ras.size(); q++) ras[q] = v[r2t[q]];
  if (isg_raster_bin_write(path, (int)ras.size(), (const double*)ras.data()) != ISG_OK)
    return fail(c, ISG_E_ARG, std::string("cannot write ") + path);
  return ISG_OK;
}

// The reference's commented-out restart path (mom.cpp:230-235): contrast read
// from a raster cg_<k>.bin as the current object.
int isg_dbim_read_cg(isg_context* c, const char* path) {
  if (!check(c)) return ISG_E_STATE;
  if (!c->dbim) return fail(c, ISG_E_STATE, "isg_dbim_setup has not been called");
  const isg::Plan& P = c->eng->plan();
  std::vector<isg::zd> ras(P.N), v(P.N);
  if (!path || isg_raster_bin_read(path, P.N, (double*)ras.data()) != ISG_OK)
    return fail(c, ISG_E_ARG, std::string("cannot read ") + (path ? path : "(null)"));
  for (int q = 0; q < P.N; q++) v[P.raster2tree[q]] = ras[q];
  if (!c->dbim->set(0, v)) return fail(c, ISG_E_DEVICE, c->dbim->error());
  return ISG_OK;
}

int isg_dbim_set(isg_context* c, int which, const double* in) {
  if (!check(c)) return ISG_E_STATE;
  if (!c->dbim) return fail(c, ISG_E_STATE, "isg_dbim_setup has not been called");
  if (!in) return fail(c, ISG_E_ARG, "null pointer");
  const isg::Plan& P = c->eng->plan();
  size_t n = (which == 4) ? (size_t)c->cfg.numtx * c->cfg.numrx : (which == 6 ? (size_t)c->cfg.numtx * P.N : (size_t)P.N);
  std::vector<isg::zd> v(n);
  std::memcpy(v.data(), in, n * sizeof(isg::zd));
  if (!c->dbim->set(which, v)) return fail(c, ISG_E_ARG, "bad field or device error");
  return ISG_OK;
}

}  // extern "C"
