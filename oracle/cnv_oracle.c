/*
 * cnv_oracle.c -- CPU restatement of cnavier's time step.  TEST INFRASTRUCTURE ONLY (the
 * checker): loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never
 * by the product library.  Build with -ffp-contract=off (oracle/Makefile does).
 *
 * Parity pins (what checks each function, so that a dropped term or a swapped index fails):
 *   orc_diff1/orc_diff2     -- bit-equal to the compiled reference Diff1/Diff2 (oracle/_ref,
 *                              tests/golden fixtures), polynomial exactness, Diff2 zero row sums.
 *   orc_apply_x/orc_apply_y -- bit-equal to orc_dense_kron_apply (literal Kron matvec) and to
 *                              the reference mtrxmul(kronecker(...)) fixtures.
 *   orc_poisson             -- bit-equal to reference poisson_SOR_log / poisson_log fixtures
 *                              (serial = lexicographic, OpenMP build = red-black); manufactured
 *                              solution; dense direct solve on small grids.
 *   orc_step/orc_run        -- bit-equal to the reference step fixtures; per-step (k, e) of the
 *                              reference run logs under output/logs.
 */
#include "cnv_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265359 /* include/poisson.h:9 */

double orc_beta(int nx, int ny)
{
    /* src/main.c:134 */
    return 0.5 * (2 / (1 + sin(ORC_PI / (nx + 1))) + 2 / (1 + sin(ORC_PI / (ny + 1))));
}

/* ---------------------------------------------------------------------------------------
 * Coefficient tables.  Each entry is num/den and is evaluated as ((double)num / den) / hp,
 * which is exactly how the reference spells every coefficient: "(double)-1 / 60 / dx" is
 * ((-1.0 / 60) / dx), "-0.5 / dx" is ((-0.5 / 1) / dx) and "2. / (dx * dx)" is
 * ((2.0 / 1) / (dx * dx)).  src/finitediff.c:57-147 (Diff1), 183-286 (Diff2).
 * ------------------------------------------------------------------------------------- */
typedef struct { double num, den; } frac;

/* one stencil row: first column offset relative to the row index (interior) or absolute
 * (left boundary rows), width, and coefficients in ascending column order. */
typedef struct { int width; frac c[7]; } stencil;

/* Diff1 */
static const stencil D1_row0      = {2, {{-1, 1}, {1, 1}}};                          /* 60-61, 79-80, 109-110 */
static const stencil D1_row1      = {3, {{-0.5, 1}, {0, 1}, {0.5, 1}}};              /* 83-85, 112-114 */
static const stencil D1_o4        = {5, {{1, 12}, {-2, 3}, {0, 1}, {2, 3}, {-1, 12}}}; /* 91-95, 116-120 */
static const stencil D1_o6        = {7, {{-1, 60}, {3, 20}, {-3, 4}, {0, 1}, {3, 4}, {-3, 20}, {1, 60}}}; /* 126-132 */
/* Diff2 */
static const stencil D2_row0      = {4, {{2, 1}, {-5, 1}, {4, 1}, {-1, 1}}};         /* 187-190, 210-213, 244-247 */
static const stencil D2_o2        = {3, {{1, 1}, {-2, 1}, {1, 1}}};                  /* 195-197, 216-218, 249-251 */
static const stencil D2_o4        = {5, {{-1, 12}, {4, 3}, {-5, 2}, {4, 3}, {-1, 12}}}; /* 224-228, 253-257 */
static const stencil D2_o6        = {7, {{1, 90}, {-3, 20}, {3, 2}, {-49, 18}, {3, 2}, {-3, 20}, {1, 90}}}; /* 263-269 */

static double coef(frac f, double hp) { return (f.num / f.den) / hp; }

/* write stencil s into row r starting at column c0 */
static void put_row(double *D, int n, int r, int c0, const stencil *s, double hp)
{
    for (int m = 0; m < s->width; m++) D[(size_t)r * n + c0 + m] = coef(s->c[m], hp);
}

int orc_diff1(int n, int o, double h, double *D)
{
    if (o != 2 && o != 4 && o != 6) return ORC_BAD_ORDER;
    if (n < o + 2) return ORC_BAD_ARG;
    memset(D, 0, sizeof(double) * (size_t)n * n);
    /* left boundary rows (one-sided, then lower-order centred) */
    const stencil *left[3] = {&D1_row0, &D1_row1, &D1_o4};
    int nleft = (o == 2) ? 1 : (o == 4) ? 2 : 3;
    const stencil *inner = (o == 2) ? &D1_row1 : (o == 4) ? &D1_o4 : &D1_o6;
    for (int r = 0; r < nleft; r++) put_row(D, n, r, 0, left[r], h);
    int half = inner->width / 2;
    for (int i = nleft; i < n - nleft; i++) put_row(D, n, i, i - half, inner, h);
    /* right boundary rows: translated copies of the left rows (src/finitediff.c:72-73,
     * 99-103, 136-145): row n-1-r holds row r's coefficients, same ascending order, ending
     * at column n-1. */
    for (int r = 0; r < nleft; r++) {
        int w = left[r]->width;
        for (int m = 0; m < w; m++)
            D[(size_t)(n - 1 - r) * n + (n - w) + m] = D[(size_t)r * n + m];
    }
    return ORC_OK;
}

int orc_diff2(int n, int o, double h, double *D)
{
    if (o != 2 && o != 4 && o != 6) return ORC_BAD_ORDER;
    if (n < o + 2) return ORC_BAD_ARG;
    memset(D, 0, sizeof(double) * (size_t)n * n);
    double h2 = h * h;
    const stencil *left[3] = {&D2_row0, &D2_o2, &D2_o4};
    int nleft = (o == 2) ? 1 : (o == 4) ? 2 : 3;
    const stencil *inner = (o == 2) ? &D2_o2 : (o == 4) ? &D2_o4 : &D2_o6;
    for (int r = 0; r < nleft; r++) put_row(D, n, r, 0, left[r], h2);
    int half = inner->width / 2;
    for (int i = nleft; i < n - nleft; i++) put_row(D, n, i, i - half, inner, h2);
    /* right boundary rows: reflections (src/finitediff.c:201-204, 232-238, 273-284):
     * D[n-1-r][n-1-c] = D[r][c]. */
    for (int r = 0; r < nleft; r++) {
        int w = left[r]->width;
        for (int c = 0; c < w; c++)
            D[(size_t)(n - 1 - r) * n + (n - 1 - c)] = D[(size_t)r * n + c];
    }
    return ORC_OK;
}

/* ---------------------------------------------------------------------------------------
 * Kronecker matvecs as stencils.  (I (x) d)[(i,j),(i',l)] = d[j][l] if i'==i (main.c:149),
 * (d (x) I)[(i,j),(k,j')] = d[i][k] if j'==j (main.c:150).  mtrxmul's k loop ascends the
 * flattened index, i.e. ascending l for DX and ascending k for DY.
 * ------------------------------------------------------------------------------------- */
void orc_apply_x(int n, const double *D, const double *f, double *out)
{
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double s = 0.0;
            const double *row = D + (size_t)j * n;
            int lo = j - 4 < 0 ? 0 : j - 4, hi = j + 4 > n - 1 ? n - 1 : j + 4;
            for (int l = lo; l <= hi; l++) /* every nonzero of a Diff row lies within +-4 */
                if (row[l] != 0.0) s += row[l] * f[(size_t)i * n + l];
            out[(size_t)i * n + j] = s;
        }
}

void orc_apply_y(int n, const double *D, const double *f, double *out)
{
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double s = 0.0;
            const double *row = D + (size_t)i * n;
            int lo = i - 4 < 0 ? 0 : i - 4, hi = i + 4 > n - 1 ? n - 1 : i + 4;
            for (int k = lo; k <= hi; k++)
                if (row[k] != 0.0) s += row[k] * f[(size_t)k * n + j];
            out[(size_t)i * n + j] = s;
        }
}

int orc_dense_kron_apply(int n, const double *D, int kron_x, const double *f, double *out)
{
    size_t N = (size_t)n * n;
    if (N > 4096) return ORC_BAD_ARG; /* (n*n)^2 terms: keep it tiny */
    for (size_t r = 0; r < N; r++) {
        double s = 0.0;
        for (size_t c = 0; c < N; c++) {
            /* kronecker(A, B)[r][c] = A[r/n][c/n] * B[r%n][c%n]  (src/linearalg.c:367) */
            double a, b;
            if (kron_x) { a = (r / n == c / n) ? 1.0 : 0.0; b = D[(r % n) * n + (c % n)]; }
            else { a = D[(r / n) * n + (c / n)]; b = (r % n == c % n) ? 1.0 : 0.0; }
            s += (a * b) * f[c];
        }
        out[r] = s;
    }
    return ORC_OK;
}

/* src/poisson.c:34-60: e = sum_{row-major} sqrt(pow(b - a, 2)). */
double orc_error(int n, const double *a, const double *b)
{
    double e = 0;
    size_t N = (size_t)n * n;
    for (size_t q = 0; q < N; q++) e += sqrt(pow(b[q] - a[q], 2));
    return e;
}

static double orc_error_threads(int n, const double *a, const double *b, int threads)
{
    if (threads <= 1) return orc_error(n, a, b);
    double e = 0;
#pragma omp parallel for reduction(+ : e) schedule(static) num_threads(threads)
    for (int i = 0; i < n; i++) {
        double r = 0;
        for (int j = 0; j < n; j++) r += fabs(b[(size_t)i * n + j] - a[(size_t)i * n + j]);
        e += r;
    }
    return e;
}

/* ---------------------------------------------------------------------------------------
 * Poisson.  SOR point update (src/poisson.c:268):
 *   u = beta * (dy*dy*(N+S) + dx*dx*(E+W) - dx*dx*dy*dy*f) / (2*(dx*dx+dy*dy)) + (1-beta)*u0
 * Gauss-Seidel (src/poisson.c:205): the same without beta and without the u0 term.
 * N = u[i+1][j], S = u[i-1][j], E = u[i][j+1], W = u[i][j-1]; u0[i][j] is the value before
 * this sweep, which in both orderings is the value the point holds when it is updated.
 * ------------------------------------------------------------------------------------- */
static inline double sor_point(const double *u, const double *f, size_t q, int n, double dx,
                               double dy, double beta, int gs)
{
    double up = u[q + n], dn = u[q - n], rt = u[q + 1], lt = u[q - 1];
    if (gs)
        return (dy * dy * (up + dn) + dx * dx * (rt + lt) - dx * dx * dy * dy * f[q]) /
               (2 * (dx * dx + dy * dy));
    return beta * (dy * dy * (up + dn) + dx * dx * (rt + lt) - dx * dx * dy * dy * f[q]) /
               (2 * (dx * dx + dy * dy)) +
           (1 - beta) * u[q];
}

int orc_poisson(const orc_params *p, const double *f, double *u, orc_stats *st)
{
    int n = p->n;
    size_t N = (size_t)n * n;
    int gs = (p->poisson_type == 1);
    if (p->poisson_type != 1 && p->poisson_type != 2) {
        st->status = ORC_BAD_TYPE;
        return ORC_BAD_TYPE;
    }
    double *u0 = (double *)malloc(sizeof(double) * N);
    if (!u0) return ORC_BAD_ARG;
    memset(u, 0, sizeof(double) * N); /* initm: psi restarts from zero (src/poisson.c:229) */
    int threads = p->threads > 0 ? p->threads : 1;
    st->status = ORC_NOT_CONVERGED;
    st->k = -1;
    st->sweeps = 0;
    st->e = 0;
    for (int k = 0; k < p->itmax; k++) {
        memcpy(u0, u, sizeof(double) * N); /* mtrxcpy(u0, u) */
        if (p->ordering == ORC_LEX) {
            for (int i = 1; i < n - 1; i++)
                for (int j = 1; j < n - 1; j++) {
                    size_t q = (size_t)i * n + j;
                    u[q] = sor_point(u, f, q, n, p->dx, p->dy, p->beta, gs);
                }
        } else {
            for (int color = 0; color < 2; color++) {
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
                for (int i = 1; i < n - 1; i++)
                    for (int j = 1 + ((i + 1 + color) & 1); j < n - 1; j += 2) {
                        size_t q = (size_t)i * n + j;
                        u[q] = sor_point(u, f, q, n, p->dx, p->dy, p->beta, gs);
                    }
            }
        }
        double e = orc_error_threads(n, u, u0, threads);
        st->sweeps = k + 1;
        st->e = e;
        if (e < p->tol) {
            st->status = ORC_OK;
            st->k = k;
            break;
        }
    }
    free(u0);
    return st->status;
}

/* ---------------------------------------------------------------------------------------
 * Step pieces.
 * ------------------------------------------------------------------------------------- */
void orc_init_state(const orc_params *p, double *w, double *psi, double *u, double *v)
{
    int n = p->n;
    size_t N = (size_t)n * n;
    memset(w, 0, sizeof(double) * N);
    memset(psi, 0, sizeof(double) * N);
    memset(u, 0, sizeof(double) * N);
    memset(v, 0, sizeof(double) * N);
    for (int i = 1; i < n - 1; i++)
        for (int j = 1; j < n - 1; j++) {
            u[(size_t)i * n + j] = p->ui; /* src/main.c:214-221 */
            v[(size_t)i * n + j] = p->vi;
        }
}

void orc_apply_velocity_bc(const orc_params *p, double *u, double *v)
{
    int n = p->n;
    /* rows first, then columns: the columns win at the corners (src/main.c:283-296) */
    for (int j = 0; j < n; j++) {
        v[j] = p->v3;
        v[(size_t)(n - 1) * n + j] = p->v4;
        u[j] = p->u3;
        u[(size_t)(n - 1) * n + j] = p->u4;
    }
    for (int i = 0; i < n; i++) {
        v[(size_t)i * n] = p->v1;
        v[(size_t)i * n + n - 1] = p->v2;
        u[(size_t)i * n] = p->u1;
        u[(size_t)i * n + n - 1] = p->u2;
    }
}

int orc_wall_vorticity(const orc_params *p, const double *u, const double *v, double *w)
{
    int n = p->n;
    size_t N = (size_t)n * n;
    double *dx1 = malloc(sizeof(double) * n * n), *dy1 = malloc(sizeof(double) * n * n);
    double *dudy = malloc(sizeof(double) * N), *dvdx = malloc(sizeof(double) * N);
    int rc = orc_diff1(n, p->order, p->dx, dx1);
    if (rc == ORC_OK) rc = orc_diff1(n, p->order, p->dy, dy1);
    if (rc == ORC_OK) {
        orc_apply_y(n, dy1, u, dudy); /* DY * u  (src/main.c:300) */
        orc_apply_x(n, dx1, v, dvdx); /* DX * v  (src/main.c:301) */
        for (int j = 0; j < n; j++) { /* src/main.c:311-320 */
            w[j] = dvdx[j] - dudy[j];
            size_t q = (size_t)(n - 1) * n + j;
            w[q] = dvdx[q] - dudy[q];
        }
        for (int i = 0; i < n; i++) {
            size_t q = (size_t)i * n;
            w[q] = dvdx[q] - dudy[q];
            q += n - 1;
            w[q] = dvdx[q] - dudy[q];
        }
    }
    free(dx1); free(dy1); free(dudy); free(dvdx);
    return rc;
}

void orc_euler(int n, double *w, const double *dwdx, const double *dwdy, const double *d2wdx2,
               const double *d2wdy2, const double *u, const double *v, double Re, double dt)
{
    size_t N = (size_t)n * n;
    for (size_t q = 0; q < N; q++) /* src/fluiddyn.c:98, every point including the edges */
        w[q] = (-u[q] * dwdx[q] - v[q] * dwdy[q] + (1. / Re) * (d2wdx2[q] + d2wdy2[q])) * dt + w[q];
}

int orc_transport(const orc_params *p, double *w, const double *u, const double *v)
{
    int n = p->n;
    size_t N = (size_t)n * n;
    double *d1x = malloc(sizeof(double) * N), *d1y = malloc(sizeof(double) * N);
    double *d2x = malloc(sizeof(double) * N), *d2y = malloc(sizeof(double) * N);
    double *a = malloc(sizeof(double) * N), *b = malloc(sizeof(double) * N);
    double *c = malloc(sizeof(double) * N), *d = malloc(sizeof(double) * N);
    int rc = orc_diff1(n, p->order, p->dx, d1x);
    if (rc == ORC_OK) rc = orc_diff1(n, p->order, p->dy, d1y);
    if (rc == ORC_OK) rc = orc_diff2(n, p->order, p->dx, d2x);
    if (rc == ORC_OK) rc = orc_diff2(n, p->order, p->dy, d2y);
    if (rc == ORC_OK) {
        orc_apply_x(n, d1x, w, a); /* dwdx   = DX  w (src/main.c:324) */
        orc_apply_y(n, d1y, w, b); /* dwdy   = DY  w (src/main.c:325) */
        orc_apply_x(n, d2x, w, c); /* d2wdx2 = DX2 w (src/main.c:333) */
        orc_apply_y(n, d2y, w, d); /* d2wdy2 = DY2 w (src/main.c:334) */
        orc_euler(n, w, a, b, c, d, u, v, p->Re, p->dt);
    }
    free(d1x); free(d1y); free(d2x); free(d2y); free(a); free(b); free(c); free(d);
    return rc;
}

int orc_velocity(const orc_params *p, const double *psi, double *u, double *v,
                 double *cont_max, double *cont_min)
{
    int n = p->n;
    size_t N = (size_t)n * n;
    double *d1x = malloc(sizeof(double) * N), *d1y = malloc(sizeof(double) * N);
    double *a = malloc(sizeof(double) * N), *b = malloc(sizeof(double) * N);
    int rc = orc_diff1(n, p->order, p->dx, d1x);
    if (rc == ORC_OK) rc = orc_diff1(n, p->order, p->dy, d1y);
    if (rc == ORC_OK) {
        orc_apply_y(n, d1y, psi, u);          /* u = DY psi (src/main.c:368,381) */
        orc_apply_x(n, d1x, psi, v);          /* v = -(DX psi) (src/main.c:367,382-383) */
        for (size_t q = 0; q < N; q++) v[q] = -v[q];
        orc_apply_x(n, d1x, u, a);            /* dudx = DX u (src/main.c:390) */
        orc_apply_y(n, d1y, v, b);            /* dvdy = DY v (src/main.c:391) */
        double mx = -DBL_MAX, mn = DBL_MAX;   /* maxel/minel (src/linearalg.c:475-545) */
        for (size_t q = 0; q < N; q++) {
            double cval = a[q] + b[q];        /* continuity (src/fluiddyn.c:149) */
            if (cval > mx) mx = cval;
            if (cval < mn) mn = cval;
        }
        if (cont_max) *cont_max = mx;
        if (cont_min) *cont_min = mn;
    }
    free(d1x); free(d1y); free(a); free(b);
    return rc;
}

int orc_step(const orc_params *p, double *w, double *psi, double *u, double *v, orc_stats *st)
{
    int n = p->n;
    size_t N = (size_t)n * n;
    orc_apply_velocity_bc(p, u, v);
    int rc = orc_wall_vorticity(p, u, v, w);
    if (rc) return rc;
    rc = orc_transport(p, w, u, v);
    if (rc) return rc;
    /* invsig(w); psi = poisson(w); invsig(w)  (src/main.c:348-363) */
    double *f = malloc(sizeof(double) * N);
    for (size_t q = 0; q < N; q++) f[q] = -w[q];
    rc = orc_poisson(p, f, psi, st);
    free(f);
    if (rc) return rc;
    return orc_velocity(p, psi, u, v, &st->cont_max, &st->cont_min);
}

int orc_run(const orc_params *p, int nsteps, double *w, double *psi, double *u, double *v,
            orc_stats *per_step)
{
    for (int t = 0; t < nsteps; t++) {
        int rc = orc_step(p, w, psi, u, v, &per_step[t]);
        if (rc) return rc;
    }
    return ORC_OK;
}
