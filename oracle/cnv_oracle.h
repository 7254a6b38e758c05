/*
 * cnv_oracle.h -- CPU restatement of cnavier's vorticity-streamfunction time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  Nothing in the
 * product library (libcnv) links it or falls back to it.
 *
 * Every function restates the reference algorithm with the same summation order, the same
 * coefficient expressions and the same operation association, so that (compiled with
 * -ffp-contract=off) it reproduces the reference binary64 results bit for bit:
 *   - derivative operators: src/finitediff.c:51-292 (Diff1/Diff2 tables), applied as the
 *     Kronecker matvecs of src/main.c:142-152 / src/linearalg.c:236-285,353-408.  A dense
 *     matvec that starts from 0.0 and adds A[r][k]*x[k] for ascending k equals the ascending
 *     sum over the nonzero coefficients only (adding a +-0 product never changes the sum).
 *   - boundary conditions and wall vorticity: src/main.c:283-320
 *   - explicit Euler: src/fluiddyn.c:71-102
 *   - Poisson solvers: src/poisson.c:34-60 (error), 176-222 (Gauss-Seidel), 224-285 (SOR),
 *     lexicographic (serial build) and red-black (OpenMP build) orderings
 *   - velocities and continuity diagnostic: src/main.c:366-395, src/fluiddyn.c:126-154,
 *     src/linearalg.c:475-545
 *
 * Arrays are n*n binary64, row-major, element (i, j) at i*n + j; row i is y, column j is x.
 */
#ifndef CNV_ORACLE_H
#define CNV_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_LEX = 0, ORC_REDBLACK = 1 };
enum { ORC_OK = 0, ORC_NOT_CONVERGED = 1, ORC_BAD_ORDER = 2, ORC_BAD_TYPE = 3, ORC_BAD_ARG = 4 };

typedef struct {
    int n;              /* nx == ny (src/main.c:142-152 requires square) */
    int order;          /* 2, 4 or 6 */
    int poisson_type;   /* 1 = Gauss-Seidel, 2 = SOR (src/main.c:349-361) */
    int ordering;       /* ORC_LEX or ORC_REDBLACK */
    int itmax;          /* poisson_max_it */
    int threads;        /* OpenMP threads for the sweeps; 1 = exact serial error() order */
    double dx, dy;      /* Lx/nx, Ly/ny (src/main.c:138-139) */
    double Re, dt, beta, tol;
    double ui, vi, u1, u2, u3, u4, v1, v2, v3, v4;
} orc_params;

typedef struct {
    int status;         /* ORC_OK or ORC_NOT_CONVERGED */
    int k;              /* logged iteration index = sweeps - 1 (src/poisson.c:275) */
    int sweeps;         /* sweeps performed */
    double e;           /* L1 norm of the last sweep's update (src/poisson.c:34-60) */
    double cont_max, cont_min;  /* max/min of Dx u + Dy v (src/main.c:387-408) */
} orc_stats;

/* Dense n x n operators, same expressions and assignment order as src/finitediff.c. */
int orc_diff1(int n, int o, double h, double *D);
int orc_diff2(int n, int o, double h, double *D);

/* out = (I (x) d) f  i.e. along j (x);   out = (d (x) I) f  i.e. along i (y). */
void orc_apply_x(int n, const double *D, const double *f, double *out);
void orc_apply_y(int n, const double *D, const double *f, double *out);
/* Literal dense Kronecker matvec (all (n*n)^2 terms) for tiny grids; kron_x = 1 builds I(x)D. */
int orc_dense_kron_apply(int n, const double *D, int kron_x, const double *f, double *out);

double orc_error(int n, const double *a, const double *b);

/* Poisson: psi is overwritten (starts from zero, src/poisson.c:229-230); f read-only. */
int orc_poisson(const orc_params *p, const double *f, double *psi, orc_stats *st);

/* Initial condition (src/main.c:178-221) and one full time step (src/main.c:276-451). */
void orc_init_state(const orc_params *p, double *w, double *psi, double *u, double *v);
int orc_step(const orc_params *p, double *w, double *psi, double *u, double *v, orc_stats *st);
int orc_run(const orc_params *p, int nsteps, double *w, double *psi, double *u, double *v,
            orc_stats *per_step);

/* Pieces of the step, exposed for unit tests. */
void orc_apply_velocity_bc(const orc_params *p, double *u, double *v);
int orc_wall_vorticity(const orc_params *p, const double *u, const double *v, double *w);
void orc_euler(int n, double *w, const double *dwdx, const double *dwdy, const double *d2wdx2,
               const double *d2wdy2, const double *u, const double *v, double Re, double dt);
int orc_transport(const orc_params *p, double *w, const double *u, const double *v);
int orc_velocity(const orc_params *p, const double *psi, double *u, double *v,
                 double *cont_max, double *cont_min);

/* SOR parameter of src/main.c:134 with the truncated PI of include/poisson.h:9. */
double orc_beta(int nx, int ny);

#ifdef __cplusplus
}
#endif
#endif
