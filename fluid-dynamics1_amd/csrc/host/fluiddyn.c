/*
 * fluiddyn.c -- the reference's time-integration API (reference include/fluiddyn.h:8-14) on the
 * GPU: each call copies its matrices to the device, runs one elementwise kernel and copies the
 * result back.  pressure() is declared by the reference's header but never defined; so here.
 */
#include <stdio.h>
#include <stdlib.h>

#include "cnv.h"
#include "fluiddyn.h"

static int g_fluiddyn_openmp = 0;

void set_fluiddyn_openmp_config(int enabled) { g_fluiddyn_openmp = enabled; }

static double *pack(mtrx A)
{
    double *p = (double *)malloc(sizeof(double) * (size_t)A.m * A.n);
    if (!p) {
        printf("** Error: insufficient memory **");
        exit(1);
    }
    for (int i = 0; i < A.m; i++)
        for (int j = 0; j < A.n; j++) p[(size_t)i * A.n + j] = A.M[i][j];
    return p;
}

static void unpack(mtrx A, const double *p)
{
    for (int i = 0; i < A.m; i++)
        for (int j = 0; j < A.n; j++) A.M[i][j] = p[(size_t)i * A.n + j];
}

static void check(int rc)
{
    if (rc != CNV_OK) {
        printf("Error: %s\n", cnv_last_error(NULL));
        exit(1);
    }
}

void euler(mtrx w, mtrx dwdx, mtrx dwdy, mtrx d2wdx2, mtrx d2wdy2, mtrx u, mtrx v, double Re, double dt)
{
    mtrx in[7] = {w, dwdx, dwdy, d2wdx2, d2wdy2, u, v};
    double *p[7];
    for (int q = 0; q < 7; q++) p[q] = pack(in[q]);
    check(cnv_euler_host(w.m * w.n, p[0], p[1], p[2], p[3], p[4], p[5], p[6], Re, dt));
    unpack(w, p[0]);
    for (int q = 0; q < 7; q++) free(p[q]);
}

static mtrx combine(mtrx a, mtrx b, int sub)
{
    double *pa = pack(a), *pb = pack(b);
    double *po = (double *)malloc(sizeof(double) * (size_t)a.m * a.n);
    if (!po) {
        printf("** Error: insufficient memory **");
        exit(1);
    }
    check(sub ? cnv_sub_host(a.m, a.n, pa, pb, po) : cnv_add_host(a.m, a.n, pa, pb, po));
    mtrx out = initm(a.m, a.n);
    unpack(out, po);
    free(pa);
    free(pb);
    free(po);
    return out;
}

mtrx continuity(mtrx dudx, mtrx dvdy) { return combine(dudx, dvdy, 0); }

/* second argument minus the first (src/fluiddyn.c:194) */
mtrx vorticity(mtrx dvdx, mtrx dudy) { return combine(dvdx, dudy, 1); }
