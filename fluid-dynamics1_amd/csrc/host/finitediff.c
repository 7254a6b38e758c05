/*
 * finitediff.c -- Diff1/Diff2 with the reference's C API (reference include/finitediff.h:8-9).
 * The coefficients come from cnv_fd_dense (fdtables.c), which evaluates every entry with the
 * reference's own expression; the same tables feed the GPU stencils.
 */
#include <stdio.h>
#include <stdlib.h>

#include "cnv_host.h"
#include "finitediff.h"

static mtrx fd_matrix(int which, int n, int o, double h)
{
    if (o != 2 && o != 4 && o != 6) {
        printf("** Error: valid orders are 2, 4 or 6 **\n"); /* src/finitediff.c:150 */
        exit(1);
    }
    if (n < o + 2) {
        printf("** Error: %d points are too few for a finite-difference order of %d **\n", n, o);
        exit(1);
    }
    double *dense = (double *)malloc(sizeof(double) * (size_t)n * (size_t)n);
    if (!dense) {
        printf("** Error: insufficient memory **");
        exit(1);
    }
    cnv_fd_dense(which, n, o, h, dense);
    mtrx D = initm(n, n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) D.M[i][j] = dense[(size_t)i * n + j];
    free(dense);
    return D;
}

mtrx Diff1(int n, int o, double dx) { return fd_matrix(1, n, o, dx); }
mtrx Diff2(int n, int o, double dx) { return fd_matrix(2, n, o, dx); }
