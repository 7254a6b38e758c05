/*
 * poisson.c -- the reference's Poisson API (reference include/poisson.h:11-20) on the GPU.
 *
 * poisson/poisson_SOR/poisson_log/poisson_SOR_log copy f to the device, run the red-black SOR
 * (Gauss-Seidel when there is no relaxation) from psi = 0 until a sweep's L1 update falls below
 * tol, and return psi as a new mtrx.  Messages and exit behaviour are the reference's
 * (src/poisson.c:99-108, 163-172, 212-221, 275-284).  One device context per grid size is kept
 * for the process and reused.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "cnv.h"
#include "poisson.h"

static int g_poisson_openmp = 0;
static cnv_ctx *g_ctx = NULL;
static int g_ctx_n = 0;

void set_poisson_openmp_config(int enabled) { g_poisson_openmp = enabled; }

/* src/poisson.c:34-60: sum over rows, then columns, of sqrt(pow(u2 - u1, 2)). */
double error(mtrx u1, mtrx u2)
{
    double e = 0;
    for (int i = 0; i < u1.m; i++)
        for (int j = 0; j < u1.n; j++) e += sqrt(pow(u2.M[i][j] - u1.M[i][j], 2));
    return e;
}

static void release_ctx(void)
{
    if (g_ctx) cnv_destroy(g_ctx);
    g_ctx = NULL;
}

static mtrx solve(mtrx f, double dx, double dy, int itmax, double tol, double beta, int type, FILE *out)
{
    const int n = f.m;
    if (f.m != f.n || n < 8) {
        printf("Error: the GPU Poisson solver needs a square grid of at least 8 x 8 points (got %d x %d)\n", f.m, f.n);
        exit(1);
    }
    if (!g_ctx || g_ctx_n != n) {
        release_ctx();
        cnv_params p;
        cnv_default_params(&p, n);
        p.diagnostics = 0;
        if (cnv_create(&p, &g_ctx) != CNV_OK) {
            printf("Error: %s\n", cnv_last_error(NULL));
            exit(1);
        }
        g_ctx_n = n;
        atexit(release_ctx);
    }
    if (cnv_set_poisson_params(g_ctx, dx, dy, beta, tol, itmax, type) != CNV_OK) {
        printf("Error: %s\n", cnv_last_error(g_ctx));
        exit(1);
    }
    double *flat = (double *)malloc(sizeof(double) * (size_t)n * n * 2);
    if (!flat) {
        printf("** Error: insufficient memory **");
        exit(1);
    }
    double *psi = flat + (size_t)n * n;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) flat[(size_t)i * n + j] = f.M[i][j];
    int k = 0;
    double e = 0;
    const int rc = cnv_poisson_host(g_ctx, flat, psi, &k, &e);
    /* a non-finite update ends the GPU solve early (CNV_DIVERGED); the reference's loop would
       spend itmax sweeps on it and then fail the same way (src/poisson.c:280-284) */
    if (rc == CNV_NOT_CONVERGED || rc == CNV_DIVERGED) {
        if (out) {
            fprintf(out, "Error: maximum number of iterations achieved for Poisson equation.\n");
            fflush(out);
        }
        free(flat);
        exit(1);
    }
    if (rc != CNV_OK) {
        printf("Error: %s\n", cnv_last_error(g_ctx));
        exit(1);
    }
    if (out) {
        fprintf(out, "Poisson equation solved with %d iterations - root-sum-of-squares error: %E\n", k, e);
        fflush(out);
    }
    mtrx u = initm(n, n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) u.M[i][j] = psi[(size_t)i * n + j];
    free(flat);
    return u;
}

mtrx poisson(mtrx f, double dx, double dy, int itmax, double tol)
{
    return solve(f, dx, dy, itmax, tol, 1.0, 1, stdout);
}

mtrx poisson_SOR(mtrx f, double dx, double dy, int itmax, double tol, double beta)
{
    return solve(f, dx, dy, itmax, tol, beta, 2, stdout);
}

mtrx poisson_log(mtrx f, double dx, double dy, int itmax, double tol, FILE *log_file)
{
    return solve(f, dx, dy, itmax, tol, 1.0, 1, log_file);
}

mtrx poisson_SOR_log(mtrx f, double dx, double dy, int itmax, double tol, double beta, FILE *log_file)
{
    return solve(f, dx, dy, itmax, tol, beta, 2, log_file);
}
