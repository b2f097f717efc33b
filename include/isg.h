/* isg.h -- C ABI of the MI355X MLFMA / BiCGSTAB / DBIM solver (libisg.so).
 *
 * Drop-in boundary for the hot path of merthidayetoglu/inverse-solver-GPU.
 * The reference exposes a C++ call surface over mutable globals (vars.h:14-26,
 * configuration in mom.cpp:4-69 and setup_mlfma.cpp:3-4).  This header keeps
 * the same operations and argument meaning with an explicit context, plain
 * pointers and sizes:
 *
 *   reference (vars.h)                          this ABI
 *   setup_mlfma() + setup_gpu() + setup_bicgs() isg_create()
 *     (setup_mlfma.cpp:50, gpu.cu:69, bicgs_gpu.cu:89)
 *   mlfma(x, r)       matvec_gpu.cu:65-79       isg_mlfma()
 *   mlfmah(x, r)      matvec_gpu.cu:80-101      isg_mlfmah()
 *   aggregate(x, r)   matvec_gpu.cu:266-368     isg_aggregate()
 *   aggregateh(x, r)  matvec_gpu.cu:369-470     isg_aggregateh()   (x is the output)
 *   bicgs(x, o, b, h) bicgs_gpu.cu:116-224      isg_bicgs()
 *   DBIM loop         mom.cpp:222-400           isg_dbim_*()
 *   incident fields / measurements  setup_born.cpp:62-65,139-168  isg_dbim_setup()
 *
 * Complex arrays are interleaved (re, im) doubles -- the memory layout of
 * std::complex<double> -- over unknowns in the reference's tree order (Morton
 * leaf boxes of 8x8 pixels, setup_mlfma.cpp:124-153); isg_geometry() returns
 * the raster-to-tree map of setup_mlfma.cpp:177-185.  Host-pointer entry points
 * copy in and out; *_dev entry points take device pointers and run on the
 * context's stream (isg_stream()).
 *
 * Every entry point returns 0 on success, a negative ISG_E* code on error
 * (isg_last_error() gives the message), and bicgs returns a positive status
 * when a system did not converge (the reference only prints, bicgs_gpu.cu:216-219).
 * One context is used by one host thread at a time.
 */
#ifndef ISG_H
#define ISG_H

#ifdef __cplusplus
extern "C" {
#endif

#define ISG_OK 0
#define ISG_E_ARG (-1)
#define ISG_E_DEVICE (-2)
#define ISG_E_STATE (-3)

#define ISG_SOLVE_CONVERGED 0
#define ISG_SOLVE_NOT_CONVERGED 1
#define ISG_SOLVE_RHO_BREAKDOWN 2
#define ISG_SOLVE_OMEGA_BREAKDOWN 3

typedef struct isg_context isg_context;

/* Configuration.  Defaults (isg_config_default) equal the reference's
 * compile-time constants: res 0.1 (mom.cpp:4), box 8 (:5), level 7 (:6),
 * toli 1e-4 (:21), iti 300 (:22), regular 0 (:32), beta 5 and ninter 30
 * (setup_mlfma.cpp:3-4); numrx/numtx/scale come from the environment in the
 * reference (mom.cpp:84-97). */
typedef struct isg_config {
  int level;       /* quadtree levels counting the root; side = 8 * 2^(level-1) pixels */
  int box;         /* pixels per leaf-box side; must be 8 */
  double res;      /* pixel size in wavelengths */
  double scale;    /* frequency scale: k0 = 2 pi scale */
  int beta;        /* excess-bandwidth parameter of the sample-count rule, in [1, 32] */
  int ninter;      /* Lagrange interpolation taps */
  int numrx;       /* far-field receivers, equispaced on 360 degrees */
  int numtx;       /* plane-wave transmitters, equispaced on 360 degrees */
  double toli;     /* BiCGSTAB relative residual */
  int iti;         /* BiCGSTAB iteration limit */
  double regular;  /* Tikhonov weight: the gradient adds regular^2 * o per transmitter */
  int max_rhs;     /* right-hand sides solved together (1 = the reference's plain loop) */
  int precision;   /* 0: FP64 operator (plain).  1: mixed -- FP32 operator tables, multipoles and
                      matrix-core arithmetic inside BiCGSTAB; FP64 vectors, dot products and
                      scalars; FP64 true-residual check with defect correction until the FP64
                      relative residual meets toli.  isg_mlfma/aggregate(h) stay FP64.
                      Refused (isg_create fails) where the translation series amplify FP32
                      rounding beyond 1e-5: beta >= 7 with the default box and resolution. */
  double mixed_margin; /* mixed precision only: the FP32 solve runs to mixed_margin * toli on its
                          recursive residual before the FP64 check, so that the check rarely
                          needs a defect-correction round; 0 < mixed_margin <= 1 (default 0.5) */
  int graph;       /* 1 (default): each BiCGSTAB iteration is captured once per solve as a HIP graph
                      and replayed; 0: launched kernel by kernel.  Results are identical. */
  double radius;   /* 0 (default): plane-wave transmitters and far-field receivers, as the reference
                      (setup_born.cpp:62-65, matvec_gpu.cu:351-361).  > 0: line-source transmitters
                      and point receivers equispaced on a circle of this radius (wavelengths):
                      incident (i/4) H0(k0 |r - rho_t|), isg_aggregate = the scattered field
                      sum_n (i/4) H0(k0 |rho_m - r_n|) res^2 x_n (dense, on the device),
                      isg_aggregateh its exact adjoint */
  int products;    /* operator of the DBIM's derivative source G (k0^2 ot tot) (mom.cpp:284-290):
                      0 (default) FP64, the plain path's matrix-vector error; 1 the FP32 operator
                      (precision 1 only).  The gradient's G^H c (mom.cpp:247-271) is FP64 either
                      way: with precision 1 it is the product that the adjoint solve's final FP64
                      residual check has already computed */
  int solve_start; /* DBIM adjoint and derivative-field solves: 0 (default) zero start, as the
                      reference (mom.cpp:286, :347); 1 the previous DBIM iteration's solution of
                      the same transmitter scaled by the residual-minimising lambda (one operator
                      application; the start is never worse than zero); the stores are dropped by
                      isg_dbim_setup* and isg_dbim_set */
  int setup;       /* operator set-up (setup_mlfma.cpp:453-631): 1 (default) the near blocks, P2M/L2P
                      patterns, shifts and translation operators are computed by device kernels in FP64;
                      0 on the host (FP64) and uploaded.  Interpolators, anterpolators and receiver
                      weights are built on the host either way. */
} isg_config;

/* Per-solve result. */
typedef struct isg_solve_info {
  int status;      /* ISG_SOLVE_* */
  int iters;       /* BiCGSTAB iterations (bicgs_gpu.cu:220) */
  double relres;   /* final recursive relative residual */
} isg_solve_info;

void isg_config_default(isg_config* cfg);
int isg_create(const isg_config* cfg, isg_context** out);
void isg_destroy(isg_context* ctx);
const char* isg_last_error(const isg_context* ctx);

/* sizes: numunk (setup: mom.cpp:137-138), numsamp per level (setup_mlfma.cpp:80-85) */
int isg_numunk(const isg_context* ctx);
int isg_level(const isg_context* ctx);
int isg_numsamp(const isg_context* ctx, int level);
double isg_k0(const isg_context* ctx);
/* pixel centres in tree order (2*numunk doubles) and raster->tree map (numunk ints); either may be NULL */
int isg_geometry(const isg_context* ctx, double* pos, int* raster2tree);
/* the HIP stream (hipStream_t) all device work of this context runs on */
void* isg_stream(const isg_context* ctx);

/* ---- reference operator surface (host arrays) ---- */
/* r = G x  (solve.cpp:25-248 / matvec_gpu.cu:65-79) */
int isg_mlfma(isg_context* ctx, const double* x, double* r);
/* r = conj(G conj x); x is not modified (the reference conjugates it in place and restores it) */
int isg_mlfmah(isg_context* ctx, const double* x, double* r);
/* r[numrx] = far-field pattern of x at the receiver angles */
int isg_aggregate(isg_context* ctx, const double* x, double* r);
/* x[numunk] = receiver back-projection of r[numrx] (output first, as the reference) */
int isg_aggregateh(isg_context* ctx, double* x, const double* r);
/* BiCGSTAB on (I - G diag(o)) x = b, or the adjoint system when herm != 0
 * (bicgs.cpp:62-87); x is the initial guess on entry and the solution on exit.
 * o is the weighted contrast o*k0^2 as passed by the reference's callers. */
int isg_bicgs(isg_context* ctx, double* x, const double* o, const double* b, int herm, isg_solve_info* info);

/* r[k] = G x[k] for nrhs host vectors ([nrhs][numunk], nrhs <= max_rhs) in one batched pass,
 * in the context's precision (the FP32 operator when precision is 1) */
int isg_mlfma_many(isg_context* ctx, const double* x, double* r, int nrhs);

/* ---- device-pointer, batched surface (vectors [nrhs][numunk], receivers [nrhs][numrx]) ---- */
int isg_mlfma_dev(isg_context* ctx, const void* x, void* r, int nrhs);
/* waits for the work queued on the context's stream (the *_dev calls return
 * before their kernels finish) */
int isg_synchronize(isg_context* ctx);
/* the same on interleaved FP32 complex vectors with the FP32 operator (precision 1
 * contexts): the matrix-vector product the mixed-precision solver runs */
int isg_mlfma_f32_dev(isg_context* ctx, const void* x, void* r, int nrhs);
int isg_aggregate_dev(isg_context* ctx, const void* x, void* r, int nrhs);
int isg_aggregateh_dev(isg_context* ctx, void* x, const void* r, int nrhs);
/* o_stride 0: one contrast shared by every right-hand side; numunk: one per right-hand side.
 * x_zero != 0 promises x == 0 on entry (skips the initial matvec). info has nrhs entries. */
int isg_bicgs_dev(isg_context* ctx, void* x, const void* o, long o_stride, const void* b, int herm, int nrhs,
                  int x_zero, isg_solve_info* info);

/* isg_bicgs_dev without allocating: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * results and errors as isg_bicgs_dev; the solver's reduction partials,
 * per-system scalars, activity flags and tolerances live in a 16 MiB static
 * device array of the library (one per GPU, kept until the process ends)
 * instead of the context's buffers, and the iterations are launched kernel by
 * kernel (no HIP graph is instantiated).  It calls no hipMalloc/hipFree.
 * Needs 48 bytes per 1024 unknowns and system plus a few hundred bytes per
 * system (2048^2 pixels x 64 systems: 12.6 MB); larger requests fail with
 * ISG_E_ARG and a message.  Every call on one GPU shares the array: calls
 * that overlap on one GPU (two host threads, two contexts) race. */
int isg_bicgs_dev_static(isg_context* ctx, void* x, const void* o, long o_stride, const void* b, int herm, int nrhs,
                         int x_zero, isg_solve_info* info);

/* multipoles (which 0) or locals (which 1) of one level after the last operator call,
 * 4^level * numsamp[level] complex values (stage checkpoints of solve.cpp:218-245) */
int isg_stage(isg_context* ctx, int which, int level, int rhs, double* out);

/* operator applications and BiCGSTAB counters since creation:
 * out[0] single-RHS matvecs (mom.cpp:61), out[1] solves, out[2] solver iterations,
 * out[3] matvecs with the FP32 operator, out[4] defect-correction solves (5 entries) */
int isg_counters(const isg_context* ctx, long long* out);

/* ---- DBIM / nonlinear-CG inversion (mom.cpp:222-400, setup_born.cpp) ---- */
typedef struct isg_dbim_stats {
  double alpha;          /* step length (mom.cpp:315) */
  double misfit;         /* sum|mesd|^2 / sum|mes|^2            (hist.txt col 1, setup_born.cpp:239) */
  double object_error;   /* sum|o - o_true|^2 / sum|o_true|^2   (col 2) */
  double object_cost;    /* sum|regular o|^2 / sum|mes|^2       (col 3) */
  double gradient_norm2; /* sum|od|^2                           (col 4) */
  double object_norm2;   /* sum|o|^2                            (col 5) */
  double beta_re, beta_im; /* Polak-Ribiere coefficient (mom.cpp:394) */
  long long solves, solve_iters, matvecs, not_converged;
  long long matvecs_fp32;  /* of matvecs, those with the FP32 operator */
  double seconds;        /* wall time of the call, device work included */
  long long field_iters; /* BiCGSTAB iterations of the total-field update phase (mom.cpp:322-336) */
  double field_seconds;  /* wall time of that phase (network and blend included) */
  double lambda_abs;     /* learned start: mean |lambda| of the blend */
  double net_seconds;    /* learned start: network inference time */
  double start_ratio;    /* learned start: mean over transmitters of the blend's initial-residual
                            ratio |r(first-order start + lambda d)| / |r(first-order start)| */
} isg_dbim_stats;

/* Incident plane waves for numtx transmitters (setup_born.cpp:62-65), synthetic
 * measurements of the true contrast o_true (tree order, numunk complex; the
 * reference's phantom file), zero initial object and the first gradient
 * (mom.cpp:222-276).  field_start: 0 = every total-field solve starts from zero
 * as in the reference (mom.cpp:324), 1 = from the previous total field, 2 = from
 * the previous total field moved to first order along the step: tot - alpha dtot,
 * dtot being the derivative field of the step-length solve (mom.cpp:284-290),
 * 3 = learned: the first-order start + lambda d with d the warm-start
 * network's correction of it and lambda minimising the initial residual per
 * transmitter, so never a worse start than the first-order one (needs
 * isg_dbim_set_warmnet before isg_dbim_iterate). */
int isg_dbim_setup(isg_context* ctx, const double* o_true, int field_start, isg_dbim_stats* st);
/* Same start from given measurements mes (numtx x numrx complex, the reference's
 * mes array after setup_born.cpp:148) instead of synthesizing them; o_true may be
 * NULL (object_error then reports sum|o|^2). */
int isg_dbim_setup_measured(isg_context* ctx, const double* mes, const double* o_true, int field_start,
                            isg_dbim_stats* st);
/* Progress of isg_dbim_setup's measurement synthesis (no counterpart in the
 * reference): fn(user, done, total) after each batch of transmitters has been
 * solved and its measurements formed, done of total transmitters; fn NULL
 * turns it off.  Kept by the context for later isg_dbim_setup calls. */
typedef void (*isg_progress_fn)(void* user, int done, int total);
int isg_set_progress(isg_context* ctx, isg_progress_fn fn, void* user);
/* One DBIM iteration (mom.cpp:277-400). */
int isg_dbim_iterate(isg_context* ctx, isg_dbim_stats* st);
/* Back to the state isg_dbim_setup* left (zero object, total fields = incident
 * fields, mesd = -mes, the first gradient and direction kept from setup rather
 * than recomputed, no previous solutions), so that the next isg_dbim_iterate
 * is iteration 1 again (mom.cpp:222-276 without its adjoint solves; no
 * counterpart in the reference, whose loop runs once: repeated timed runs of
 * the first iterations). */
int isg_dbim_restart(isg_context* ctx);
/* which: 0 o, 1 od (gradient), 2 ot (direction), 3 mes, 4 mesd, 5 inc, 6 tot,
 * 7 true contrast (get only), 8 dtot: the Frechet-derivative fields of the
 * last step-length phase, numtx x numunk (get only; field_start 2 or 3 or
 * solve_start 1).
 * Vectors are numunk (x numtx for inc/tot) complex; mes/mesd numtx x numrx.
 * isg_dbim_set accepts 0, 1, 2, 4 and 6 (checkpoint / resume). */
int isg_dbim_get(isg_context* ctx, int which, double* out);
int isg_dbim_set(isg_context* ctx, int which, const double* in);

/* ---- the reference's files (setup_born.cpp:105-121, 171-242; mom.cpp:220) ---- */
/* cg_<k>.bin (which 0: current contrast) or cg_ref.bin (which 7: true
 * contrast): numunk complex doubles in raster order (replaces the unkmap
 * gather + fwrite of write_error, setup_born.cpp:209-218). */
int isg_dbim_write_cg(isg_context* ctx, const char* path, int which);
/* contrast from a raster cg_<k>.bin (restart, mom.cpp:230-235 commented out upstream) */
int isg_dbim_read_cg(isg_context* ctx, const char* path);
/* hist.txt line "%e %e %e %e %e\n" (setup_born.cpp:239); create = 1 truncates first */
int isg_hist_append(const char* path, const isg_dbim_stats* st, int create);
/* phantom text file: numunk lines "%f", raster order; out = val * scale + 0i
 * (the reference reads with scale 0.1, setup_born.cpp:105-112) */
int isg_phantom_read_txt(const char* path, int numunk, double scale, double* raster_out);
/* writes Re(raster[n]) / scale per line (the inverse of the reader) */
int isg_phantom_write_txt(const char* path, int numunk, double scale, const double* raster_in);
/* raw raster files of numunk complex doubles */
int isg_raster_bin_write(const char* path, int numunk, const double* raster);
int isg_raster_bin_read(const char* path, int numunk, double* raster);

/* ---- autotuner (kernel variants per stage and level, measured on the device) ----
 * stage 0 near-field tile width (16, 32 or 64 boxes; 64 for the FP32
 * operator only; 128, FP32 operator only: the split-FP16 kernel, 64-box tiles
 * of two right-hand sides on the FP16 matrix cores with FP32 operands carried
 * as hi + lo, used for launches of 4 or more right-hand sides; 256, likewise:
 * 128-box tiles of one right-hand side, grids of 16 x 8 leaf boxes or more), 1 P2M tile
 * width (16, 32 or 64; 128, FP32 operator only: the split-FP16 P2M, 32-box
 * tiles of two right-hand sides, 4 or more right-hand sides), 2 M2L tile
 * variant of `level` (0: 4x4 boxes x 8 samples, 1: 8x8 x 8, 2: 8x8 x 16,
 * 3: 16x16 x 8, 4: 8x8 x 32, 5: 16x16 x 16; 6, 7, 8: 8x8 x 16, 32, 8 with the
 * operators held in registers; 9: 8x8 x 16, one thread per parent with its 36
 * sources in registers; 4, 5 and 7 FP32 only), 3 / 4 M2M / L2L
 * kernel of `level` (0 band MFMA, 1 chunked LDS, 2 direct, 3 / 4 band MFMA with
 * 16 / 32 columns per workgroup), 5 leaf fusion (`level` ignored; 1: the
 * leaf-level M2L and the L2L into the leaf level run as one kernel that
 * writes each leaf local expansion once, 0: separate); value -1 = built-in
 * rule (separate).  precision 0 = FP64 operator, 1 = FP32 operator of the mixed path;
 * batch_class 0 for fewer than 16 right-hand sides, 1 otherwise.
 * isg_autotune times every admissible variant (reps launches each) and keeps
 * the fastest for batch_class(nrhs) and every operator precision the context has.
 * The file holds lines "tree-depth precision batch_class stage level value";
 * entries for other tree depths are ignored on load. */
int isg_autotune(isg_context* ctx, int nrhs, int reps);
int isg_tuning_set(isg_context* ctx, int precision, int batch_class, int stage, int level, int value);
int isg_tuning_get(isg_context* ctx, int precision, int batch_class, int stage, int level, int* value);
int isg_tuning_save(isg_context* ctx, const char* path);
int isg_tuning_load(isg_context* ctx, const char* path);

/* ---- live kernel timing (HIP events on the context's stream; the reference
 * times whole matvecs with cudaEvents, matvec_gpu.cu:117-122) ----
 * kernel 0: near field + L2P on the matrix cores, one launch per matvec
 * kernel 1: M2L over all levels of one matvec
 * kernel 2: BiCGSTAB x/r update with its two dot products
 * kernel 3, 4: as 0, 1 for the FP32 operator of the mixed path
 * kernel 5: P2M of one matvec (either precision)
 * kernel 6: M2M and L2L interpolation of one matvec (either precision)
 * kernel 7: receiver operator (isg_aggregate / isg_aggregateh and the DBIM
 *           loop's receiver products), one launch set per call
 * isg_profile(ctx, 1) clears the totals and starts recording. */
int isg_profile(isg_context* ctx, int enable);
int isg_profile_read(isg_context* ctx, int kernel, double* total_ms, long long* launches);

/* ---- learned warm start of the total-field solves (no counterpart in the
 * reference, whose total-field solves restart from zero, mom.cpp:324) ----
 * A 4-level convolutional encoder-decoder (32 base channels, BF16 weights,
 * FP32 accumulation) mapping per transmitter 8 channels (Re/Im contrast,
 * Re/Im incident field, Re/Im first-order start tot_prev - alpha dtot, Re/Im
 * its step -alpha dtot; the field channels divided by the transmitter's
 * incident-field rms) to a 2-channel correction of the first-order start.
 * The image enters space-to-depth by 2 (the network runs at half the pixel
 * resolution with 32 input channels) and leaves depth-to-space.  Images are
 * raster order, NHWC FP32, H and W multiples of 16.
 * Calls without a context run on the null stream and return after the work. */
typedef struct isg_warmnet isg_warmnet;
/* Workspace for max_batch images of max_pixels pixels (multiple of 16 x 16);
 * He-normal weights from seed. */
int isg_warmnet_create(int max_batch, long max_pixels, unsigned long long seed, isg_warmnet** out);
void isg_warmnet_destroy(isg_warmnet* net);
const char* isg_warmnet_last_error(const isg_warmnet* net);
/* Flat FP32 parameters: per layer W[ks*ks][cout][cin] then bias[cout]. */
long isg_warmnet_num_params(void);
int isg_warmnet_get_params(isg_warmnet* net, float* out);
int isg_warmnet_set_params(isg_warmnet* net, const float* in);  /* also resets Adam */
int isg_warmnet_get_grads(isg_warmnet* net, float* out);
/* 9 scales: input channel c is multiplied by s[c] (c < 8), the output by s[8]. */
int isg_warmnet_set_scales(isg_warmnet* net, const float* s);
int isg_warmnet_get_scales(isg_warmnet* net, float* s);
/* in [batch][H][W][8], out [batch][H][W][2], device pointers */
int isg_warmnet_forward_dev(isg_warmnet* net, const void* in, void* out, int batch, int H, int W);
/* forward + backward of mean((out - target)^2) over the 2 output channels;
 * gradients kept for isg_warmnet_adam / isg_warmnet_get_grads */
int isg_warmnet_loss_grad_dev(isg_warmnet* net, const void* in, const void* target, int batch, int H, int W,
                              double* loss);
/* the same two calls on host arrays (staged through device buffers the handle keeps) */
int isg_warmnet_forward(isg_warmnet* net, const float* in, float* out, int batch, int H, int W);
int isg_warmnet_loss_grad(isg_warmnet* net, const float* in, const float* target, int batch, int H, int W,
                          double* loss);
/* Adam step (beta1 0.9, beta2 0.999, eps 1e-8) with learning rate lr */
int isg_warmnet_adam(isg_warmnet* net, float lr);
/* weight file: "ISGWNET2", uint32 count, 9 float scales, float parameters */
int isg_warmnet_save(isg_warmnet* net, const char* path);
int isg_warmnet_load(isg_warmnet* net, const char* path);
/* Network used by field_start 3 of this context's DBIM (not owned). */
int isg_dbim_set_warmnet(isg_context* ctx, isg_warmnet* net);

/* ---- seeded phantoms (raster order, 2*side*side doubles) ---- */
/* n Gaussian blobs, centres in the inner 80%, sigma U(0.3,1), amplitude U(0.1,0.5)+iU(0.01,0.05) */
int isg_phantom_blobs(int side, double res, int n, unsigned long long seed, double* out);
/* n ellipses, semi-axes U(0.02,0.08)*side*res, value U(vmin,vmax) + i*imag, later ones on top */
int isg_phantom_ellipses(int side, double res, int n, unsigned long long seed, double vmin, double vmax,
                         double imag, double* out);

#ifdef __cplusplus
}
#endif

#endif /* ISG_H */
