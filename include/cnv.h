/*
 * cnv.h -- device-resident C ABI of the MI355X (gfx950) vorticity-streamfunction step.
 *
 * This is the layer the cnavier driver uses at scale.  It has no counterpart with these names in
 * the reference: the reference's step is inline in src/main.c:276-451 and calls euler()
 * (include/fluiddyn.h:8), poisson_SOR_log()/poisson_log() (include/poisson.h:15-16) and the
 * dense mtrxmul() (include/linearalg.h:30) on host matrices.  The legacy headers in this
 * directory (poisson.h, finitediff.h, fluiddyn.h, linearalg.h, config.h, utils.h) keep those
 * exact signatures and are implemented on top of this API.
 *
 * Conventions: one context = one cavity on one GPU, one HIP stream (default priority), device
 * buffers owned by the context.  Grids are n x n binary64, row-major, element (i, j) at
 * i*n + j, row i is y and column j is x (src/main.c:283-296).  Every call returns an int status
 * (CNV_OK = 0); cnv_last_error() gives the message.  A context is not thread-safe.
 */
#ifndef CNV_H
#define CNV_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CNV_OK 0
#define CNV_ERR_ARG 1          /* bad argument (size, order, type, precision ...)           */
#define CNV_ERR_HIP 2          /* a HIP runtime call failed                                 */
#define CNV_NOT_CONVERGED 3    /* Poisson hit itmax (src/poisson.c:280-284 would exit(1))   */
#define CNV_ERR_NOMEM 4
#define CNV_ERR_NO_DEVICE 5
#define CNV_DIVERGED 6         /* a Poisson solve's L1 update became NaN or infinite, or a step's
                                  largest |u| or |v| exceeded blowup_speed (or was NaN); the
                                  reference has no such check: its solve runs to itmax and exits
                                  (src/poisson.c:280-284)                                     */

#define CNV_PREC_FP64 0        /* plain FP64, bit-faithful red-black SOR                      */
#define CNV_PREC_FP32_DC 1     /* FP32 stencils and SOR sweeps inside FP64 defect correction */

typedef struct cnv_ctx cnv_ctx;

typedef struct {
    int n;                 /* nx == ny (src/main.c:142-152 needs square), n >= 8            */
    int order;             /* finite-difference order 2, 4 or 6 (src/finitediff.c)          */
    int poisson_type;      /* 1 = Gauss-Seidel, 2 = SOR (src/main.c:349-361)                */
    int precision;         /* CNV_PREC_FP64 or CNV_PREC_FP32_DC                             */
    int itmax;             /* poisson_max_it: sweeps allowed per solve                       */
    int device;            /* HIP device ordinal                                            */
    double dx, dy;         /* Lx/nx and Ly/ny (src/main.c:138-139)                          */
    double Re, dt;
    double beta;           /* SOR relaxation, src/main.c:134 by default                      */
    double tol;            /* stop when the L1 norm of a sweep's update < tol (src/poisson.c:272-273) */
    double ui, vi;         /* initial interior velocity (src/main.c:214-221)                */
    double u1, u2, u3, u4; /* wall velocities: column 0, column n-1, row 0, row n-1          */
    double v1, v2, v3, v4;
    /* FP32_DC only */
    double inner_rtol;     /* inner FP32 solve: fold once its L1 update < inner_rtol * L1 norm of
                              the inner iterate, tested on the last sweep of a launch (every
                              inner_test_every-th launch)                                     */
    int max_outer;         /* defect-correction iterations allowed per solve                */
    /* tuning (0 = automatic) */
    int sweeps_per_launch; /* SOR sweeps fused per kernel launch (compiled: 1, 2, 4, 5, 6, 8; 0 =
                              5 or 6 with the tile shape, both chosen by grid size)          */
    int tile_waves;        /* wave64s per SOR workgroup (compiled shapes only)               */
    int wave_rows;         /* grid rows held in registers per wave (compiled shapes only)    */
    int launches_per_poll; /* SOR launches queued between host polls of the convergence flag
                              (0: per solve from the previous solve's length, 66 per captured
                              chunk with use_graph (72 for FP32_DC: chunks hold multiples of 6,
                              12 for FP32_DC), 2 for a batch)                               */
    int diagnostics;       /* 1 = continuity max/min every step (src/main.c:387-408)        */
    int use_graph;         /* 1 = each step replays captured HIP graphs (the pre-solve
                              launches, chunks of launches_per_poll SOR launches and tail
                              chunks of 6 (12 for FP32_DC), the post-solve launches; FP32_DC also the residual
                              and correction of each outer iteration, with the sweep budget
                              kept on the device); results identical to the eager launches   */
    double inner_beta;     /* FP32_DC: relaxation of the inner FP32 SOR solves (0 = beta)     */
    int warm_start;        /* 1 = each Poisson solve starts from the previous step's psi
                              instead of 0 (src/poisson.c:229 starts from 0): fewer sweeps,
                              results that differ within the tolerance; off for parity.
                              One-cavity contexts and the adjoint (whose adjoint solves then
                              start from the previous reverse step's solution); a batch
                              refuses it                                                     */
    int transport_fp32;    /* FP32_DC: 1 (default) = the transport and velocity stencils run in
                              FP32, each tap entering as an FP64 difference from the point's own
                              value (so their rounding is relative to the derivative, and the
                              refined trajectory does not drift from the FP64 one: about 2e-7
                              relative over 600 steps at 256^2, Re 1000, mostly the velocities'
                              own rounding); the state update stays FP64.  0 = FP64 stencils
                              (1e-10 there).  Ignored by FP64                                 */
    double blowup_speed;   /* > 0: a step after which some |u| or |v| exceeds this (or is NaN)
                              ends with CNV_DIVERGED (the explicit scheme has gone unstable);
                              0 = off, as in the reference.  In a batch the cavity is dropped
                              from later steps and the others go on                           */
    int inner_test_every;  /* FP32_DC: the inner FP32 solves test their stopping rule on the last
                              sweep of every k-th SOR launch only (the others skip the sums and
                              their decision); k a power of two up to 64.  0 = 4 for a context (6.1 against 6.4 us per
                              launch at 1024^2), 1 for a batch, whose outer iterations run in
                              lockstep so that a late test on any cavity's short inner solve
                              holds up all of them.  A solve may run up to k*S - 1 sweeps past
                              the first sweep that met the rule                               */
} cnv_params;

typedef struct {
    int status;            /* CNV_OK, CNV_NOT_CONVERGED or CNV_DIVERGED                      */
    int k;                 /* logged Poisson index = sweeps - 1 (src/poisson.c:275)          */
    int sweeps;            /* SOR sweeps of this step (FP64) / inner FP32 sweeps (DC)        */
    int outer;             /* defect-correction iterations (DC), 0 for FP64                  */
    double e;              /* L1 update of the last sweep (FP64) / of the last tested inner
                              FP32 sweep (DC)                                                 */
    double cont_max, cont_min; /* continuity diagnostic, if enabled                          */
} cnv_step_stats;

/* Reference defaults for an n x n cavity: load_default_config() values (src/config.c:47-87)
 * except n, with beta from src/main.c:134 and dx = 1/n. */
void cnv_default_params(cnv_params *p, int n);
/* sizeof(cnv_params) as compiled into the library: bindings check their layout against it. */
size_t cnv_params_size(void);
double cnv_reference_beta(int nx, int ny);

int cnv_create(const cnv_params *p, cnv_ctx **out);
void cnv_destroy(cnv_ctx *ctx);
const char *cnv_last_error(const cnv_ctx *ctx);   /* ctx may be NULL: this thread's last error */
const char *cnv_status_string(int status);

/* State transfer (host arrays of n*n doubles; any pointer may be NULL to skip a field).
 * A new context holds the reference initial condition (src/main.c:178-221). */
int cnv_set_state(cnv_ctx *ctx, const double *w, const double *psi, const double *u, const double *v);
int cnv_get_state(cnv_ctx *ctx, double *w, double *psi, double *u, double *v);

/* Advance nsteps time steps (src/main.c:276-451).  stats: nsteps entries or NULL.  Stops at the
 * first step whose Poisson solve does not converge and returns CNV_NOT_CONVERGED, or at the first
 * step that diverged and returns CNV_DIVERGED (the state then holds that step's non-finite or
 * blown-up fields). */
int cnv_step(cnv_ctx *ctx, int nsteps, cnv_step_stats *stats);

/* Timed variant: wall time of the step loop measured with HIP events on the context stream. */
int cnv_step_timed(cnv_ctx *ctx, int nsteps, cnv_step_stats *stats, double *elapsed_ms);

/* Output without stopping the GPU (no counterpart in the reference, which writes its VTK files
 * between steps from host memory).  cnv_snapshot_begin queues, behind the steps already queued,
 * a device copy of the state (w, psi, u, v) and its transfer to pinned host memory owned by the
 * context, on a stream of its own, and returns at once: the steps that follow run while it
 * transfers.  cnv_snapshot_wait blocks until that transfer is complete and returns host pointers
 * to n*n doubles each (any may be NULL), valid until the next cnv_snapshot_begin.  It may be
 * called from another thread while the context steps (its failures are reported in that
 * thread's cnv_last_error(NULL)); cnv_snapshot_begin and everything else stay on the context's
 * thread. */
int cnv_snapshot_begin(cnv_ctx *ctx);
int cnv_snapshot_wait(cnv_ctx *ctx, const double **w, const double **psi, const double **u,
                      const double **v);

/* Device pointers of the state (w, psi, u, v) for zero-copy callers; valid until the next step. */
int cnv_device_state(cnv_ctx *ctx, double **w, double **psi, double **u, double **v);

/* Changes the Poisson solve's parameters of an existing context (spacing used by the SOR
 * update, relaxation, tolerance, sweep budget, 1 = Gauss-Seidel / 2 = SOR).  The transport
 * stencils keep the spacing the context was created with. */
int cnv_set_poisson_params(cnv_ctx *ctx, double dx, double dy, double beta, double tol, int itmax,
                           int poisson_type);

/* Poisson solve alone on this context's grid: psi = SOR(f) from zero (src/poisson.c:224-285),
 * f given as a host array; psi written to host.  k/e as logged by the reference.  Runs in the
 * context's own buffers: the context's psi (cnv_get_state) becomes this solve's result. */
int cnv_poisson_host(cnv_ctx *ctx, const double *f, double *psi, int *k, double *e);

/* Same on device pointers (f, psi: n*n doubles in device memory), on the context stream. */
int cnv_poisson_device(cnv_ctx *ctx, const double *d_f, double *d_psi, int *k, double *e);

/* No counterpart in the reference.  For callers that cannot allocate device memory -- code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc: the Poisson solve of
 * cnv_poisson_device without a context and without allocating anything.  psi = SOR(f) from zero
 * (src/poisson.c:224-285) on an n x n grid with spacing (dx, dy), relaxation beta, tolerance tol,
 * sweep budget itmax, poisson_type 1 (Gauss-Seidel) or 2 (SOR); f and psi are device pointers to
 * n*n doubles (psi is written; f is read only).  Runs on the null stream of the current device
 * and returns when psi is complete; f must be complete when it is called (work queued on a
 * non-blocking stream is not waited for).  The solve's scratch (the second and third iterate
 * planes, the per-tile partial sums, the state slots) is a static device array of the library,
 * sized for n <= CNV_NOALLOC_MAX_N (about 17 MB, held on each device the library runs on until
 * the process ends); larger n is refused with CNV_ERR_ARG.  Calls that overlap on one device
 * share that array and race.  Results, k and e equal cnv_poisson_device's on a context created
 * with the same n, dx, dy, beta, tol, itmax, poisson_type and the default launch shape. */
#define CNV_NOALLOC_MAX_N 1024
int cnv_poisson_device_noalloc(int n, double dx, double dy, double beta, double tol, int itmax,
                               int poisson_type, const double *d_f, double *d_psi, int *k, double *e);

/* Explicit Euler update alone (src/fluiddyn.c:71-102) on host arrays of n*n doubles. */
int cnv_euler_host(int n, double *w, const double *dwdx, const double *dwdy, const double *d2wdx2,
                   const double *d2wdy2, const double *u, const double *v, double Re, double dt);

/* Continuity field Dx u + Dy v's pointwise sum (src/fluiddyn.c:126-154) and vorticity's
 * second-minus-first (src/fluiddyn.c:179-207) on host arrays: out = a + b / out = b - a. */
int cnv_add_host(int m, int n, const double *a, const double *b, double *out);
int cnv_sub_host(int m, int n, const double *a, const double *b, double *out);

/* Measurement hooks used by bench.py.
 * cnv_time_sor: replays `launches` SOR kernel launches (sweeps_per_launch sweeps each) on the
 * current f with convergence disabled and returns the mean launch duration in ms, measured with
 * HIP events on the stream the kernel runs on, after 1000 untimed launches (the steady state of
 * a solve); the timed launches are captured in a graph and replayed, as a solve's chunks are.
 * Restores nothing: call on a scratch context. */
int cnv_time_sor(cnv_ctx *ctx, int launches, double *mean_ms);
int cnv_sweeps_per_launch(const cnv_ctx *ctx);
/* Counter calibration: copies nbytes between two scratch device buffers with a plain copy kernel
 * (nbytes read + nbytes written), so that profiler byte counters can be scaled. */
int cnv_calibration_copy(cnv_ctx *ctx, long long nbytes, double *ms);
int cnv_device_name(int device, char *buf, int len);

/* ---- batched cavities ------------------------------------------------------------------------
 * Many independent cavities (own n, Re, dt, beta, tol, itmax, wall velocities) advanced together
 * on one GPU: every phase of a time step is one launch covering all cavities, and each SOR
 * launch advances every cavity whose solve is still running.  precision, order, poisson_type,
 * device and the launch shape come from items[0] and must be the same in every item.  Each
 * cavity's result equals the same cavity run alone (cnv_create / cnv_step) with the same launch
 * shape, bit for bit.  No counterpart in the reference (one cavity per process there). */
typedef struct cnv_batch cnv_batch;
int cnv_batch_create(const cnv_params *items, int count, cnv_batch **out);
void cnv_batch_destroy(cnv_batch *b);
int cnv_batch_size(const cnv_batch *b);
const char *cnv_batch_last_error(const cnv_batch *b);   /* b may be NULL: this thread's last batch error */
/* Advance every live cavity by nsteps.  stats: NULL or count*nsteps entries, stats[t*count + i].
 * A cavity that diverges (CNV_DIVERGED: a non-finite Poisson update, or its blowup_speed exceeded)
 * is reported in that step's stats, keeps its last fields and is left out of every later step
 * (stats status CNV_DIVERGED, 0 sweeps); the others go on and the call returns CNV_OK.  A live
 * cavity whose solve runs out of sweeps ends the call with CNV_NOT_CONVERGED. */
int cnv_batch_step(cnv_batch *b, int nsteps, cnv_step_stats *stats);
int cnv_batch_step_timed(cnv_batch *b, int nsteps, cnv_step_stats *stats, double *elapsed_ms);
/* State of cavity i (host arrays of n_i*n_i doubles; NULL pointers skip a field). */
int cnv_batch_get_state(cnv_batch *b, int i, double *w, double *psi, double *u, double *v);
/* Replaces cavity i's state (NULL pointers keep a field); a diverged cavity becomes live again. */
int cnv_batch_set_state(cnv_batch *b, int i, const double *w, const double *psi, const double *u,
                        const double *v);
/* Per-cavity status: status[i] = CNV_OK (live) or CNV_DIVERGED.  Returns the number of live
 * cavities, or a negative CNV_ERR_* code. */
int cnv_batch_status(const cnv_batch *b, int *status);
/* Launch-shape autotuner: times `launches` batched SOR launches of every compiled shape on a
 * synthetic source over the whole batch, keeps the fastest per sweep and reports it.  Leaves w,
 * u, v as they were; the Poisson iterates are zeroed (every step recomputes psi from zero). */
int cnv_batch_autotune(cnv_batch *b, int launches, int *sweeps_per_launch, int *tile_waves,
                       int *wave_rows, double *ms_per_sweep);

/* ---- adjoint (reverse-mode gradient) -----------------------------------------------------------
 * Objective: E = 1/2 dx dy * sum over all points of (u^2 + v^2) after nsteps FP64 steps of the
 * cavity described by p, started from the vorticity field w0 (u = ui, v = vi inside, 0 on the
 * walls) with lid velocity U (u4).  Gradients dE/dw0 (n*n, exactly 0 on the walls: step 0
 * overwrites the wall vorticity before reading it) and dE/dU by hand-written adjoint kernels of
 * the whole step (transposed advection/diffusion stencils, the self-adjoint Poisson solve, the
 * adjoints of the velocity and wall-vorticity boundary conditions), with the forward state
 * checkpointed every checkpoint_every steps in device memory and recomputed per segment.
 * Every Poisson solve of the adjoint problem (forward, recompute, adjoint) stops when its L1
 * update falls below rtol times the L1 norm of the iterate (rtol_forward / rtol_adjoint), so
 * that the gradient is that of a fully converged trajectory.  FP64, SOR (poisson_type 2), n >= 10.
 * No counterpart in the reference. */
typedef struct cnv_adjoint cnv_adjoint;
int cnv_adj_create(const cnv_params *p, int checkpoint_every, double rtol_forward, double rtol_adjoint,
                   cnv_adjoint **out);
void cnv_adj_destroy(cnv_adjoint *a);
const char *cnv_adj_last_error(const cnv_adjoint *a);   /* a may be NULL: this thread's last adjoint error */
/* Forward run from (w0, U) for nsteps; stores the checkpoints; returns E. */
int cnv_adj_forward(cnv_adjoint *a, const double *w0, double U, int nsteps, double *energy);
/* Gradient of the last forward run's E: dE_dw0 (host, n*n) and dE_dU (either may be NULL); repeated
 * calls after one forward run return the same gradient; refused after a forward run that failed.
 * A second host thread recomputes the segments on a stream of its own while this thread runs the
 * reverse steps of the segment after (two segment buffers of checkpoint_every x 3 planes; the
 * forward run leaves its last segment's states in one of them); the result is the same as running
 * them one after the other.  Returns when both are done. */
int cnv_adj_backward(cnv_adjoint *a, double *dE_dw0, double *dE_dU);
/* SOR sweeps of the last forward run and of the last backward run's adjoint solves (its
 * recompute not counted). */
int cnv_adj_sweeps(const cnv_adjoint *a, long long *forward, long long *adjoint);

#ifdef __cplusplus
}
#endif
#endif
