/* poisson.h -- Poisson solvers for the streamfunction, same C API as the reference
 * (reference include/poisson.h:9-20), solved on the GPU.
 *
 * Each solver starts from psi = 0, applies the reference point update (Gauss-Seidel, or SOR with
 * relaxation beta) in red-black order on one MI355X, and stops at the first sweep whose L1
 * update norm is below tol; it returns a newly allocated psi (free it with freem).  The logged
 * index is sweeps - 1.  Non-convergence within itmax sweeps prints the reference's message and
 * calls exit(1).  Grids must be square with n >= 8.  The red-black order is the reference's
 * OpenMP build order (src/poisson.c:237-262); its serial build sweeps lexicographically and so
 * stops at a different sweep of the same fixed point. */
#ifndef POISSON_H_INCLUDED
#define POISSON_H_INCLUDED

#include <stdio.h>

#include "linearalg.h"

#define PI 3.14159265359

double error(mtrx u1, mtrx u2); /* sum of |u2 - u1| in row-major order (src/poisson.c:34-60) */
/* Gauss-Seidel (src/poisson.c:62) */
mtrx poisson(mtrx f, double dx, double dy, int itmax, double tol);
/* SOR with relaxation beta (src/poisson.c:111) */
mtrx poisson_SOR(mtrx f, double dx, double dy, int itmax, double tol, double beta);
/* Gauss-Seidel logging its sweep count to log_file, which may be NULL (src/poisson.c:176) */
mtrx poisson_log(mtrx f, double dx, double dy, int itmax, double tol, FILE *log_file);
/* SOR logging its sweep count (src/poisson.c:224) */
mtrx poisson_SOR_log(mtrx f, double dx, double dy, int itmax, double tol, double beta, FILE *log_file);
/* accepted for compatibility; the GPU path is used (src/poisson.c:15) */
void set_poisson_openmp_config(int enabled);

#endif
