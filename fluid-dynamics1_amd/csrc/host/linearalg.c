/*
 * linearalg.c -- the reference's mtrx container API (reference include/linearalg.h:6-43,
 * src/linearalg.c), host side.  Matrices are row-pointer arrays with one allocation per row, so
 * mtrx values returned by this library are released with freem() exactly like the reference's.
 * mtrxmul keeps the reference's arithmetic: C[i][j] starts from 0.0 and accumulates A[i][k]*B[k][j]
 * for ascending k (src/linearalg.c:269-280).
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "linearalg.h"

static int g_la_openmp = 0;

void set_openmp_config(int enabled) { g_la_openmp = enabled; }

static void fatal(const char *msg)
{
    printf("%s", msg);
    exit(1);
}

double **allocm(int m, int n)
{
    if (m < 1 || n < 1) fatal("** Error: invalid parameter **\n");
    double **rows = (double **)malloc((size_t)m * sizeof(double *));
    if (!rows) fatal("** Error: insufficient memory **");
    for (int i = 0; i < m; i++) {
        rows[i] = (double *)malloc((size_t)n * sizeof(double));
        if (!rows[i]) fatal("** Error: insufficient memory **");
    }
    return rows;
}

double **freem(mtrx A)
{
    if (A.M == NULL) return NULL;
    if (A.m < 1 || A.n < 1) fatal("** Error: invalid parameter **\n");
    for (int i = 0; i < A.m; i++) free(A.M[i]);
    free(A.M);
    return NULL;
}

void zerosm(mtrx A)
{
    for (int i = 0; i < A.m; i++)
        for (int j = 0; j < A.n; j++) A.M[i][j] = 0;
}

mtrx initm(int m, int n)
{
    mtrx A;
    A.M = allocm(m, n);
    A.m = m;
    A.n = n;
    zerosm(A);
    return A;
}

mtrx eye(int n)
{
    mtrx A = initm(n, n);
    for (int i = 0; i < n; i++) A.M[i][i] = 1;
    return A;
}

void invsig(mtrx A)
{
    for (int i = 0; i < A.m; i++)
        for (int j = 0; j < A.n; j++) A.M[i][j] = -A.M[i][j];
}

void mtrxcpy(mtrx A, mtrx B)
{
    for (int i = 0; i < A.m; i++)
        for (int j = 0; j < A.n; j++) A.M[i][j] = B.M[i][j];
}

double maxel(mtrx A)
{
    double best = -DBL_MAX;
    for (int i = 0; i < A.m; i++)
        for (int j = 0; j < A.n; j++)
            if (A.M[i][j] > best) best = A.M[i][j];
    return best;
}

double minel(mtrx A)
{
    double best = DBL_MAX;
    for (int i = 0; i < A.m; i++)
        for (int j = 0; j < A.n; j++)
            if (A.M[i][j] < best) best = A.M[i][j];
    return best;
}

mtrx mtrxmul(mtrx A, mtrx B)
{
    if (A.n != B.m) {
        printf("** Error: the first matrix number of columns must be equal to the second matrix number of rows **\n");
        printf("Columns of first matrix: %d\n", A.n);
        printf("Rows of second matrix: %d\n", B.m);
        exit(1);
    }
    mtrx C = initm(A.m, B.n);  /* every C[i][j] starts at +0.0 */
    /* i-k-j order: each C[i][j] still accumulates its products for ascending k */
    for (int i = 0; i < A.m; i++) {
        double *c = C.M[i];
        for (int k = 0; k < A.n; k++) {
            const double a = A.M[i][k];
            const double *b = B.M[k];
            for (int j = 0; j < B.n; j++) c[j] += a * b[j];
        }
    }
    return C;
}

mtrx kronecker(mtrx A, mtrx B)
{
    /* (A (x) B)[i][j] = A[i / B.m][j / B.n] * B[i % B.m][j % B.n].  The reference divides by
     * A.n in both places (src/linearalg.c:367), which is the same for the square, equal-sized
     * factors it is used with. */
    mtrx C;
    C.m = A.m * B.m;
    C.n = A.n * B.n;
    C.M = allocm(C.m, C.n);
    for (int i = 0; i < C.m; i++)
        for (int j = 0; j < C.n; j++) C.M[i][j] = A.M[i / B.m][j / B.n] * B.M[i % B.m][j % B.n];
    return C;
}

mtrx reshape(mtrx A, int m, int n)
{
    if (A.m * A.n != m * n) {
        printf("** Error: the reshaped matrix must have the same number of elements **\n");
        printf("Number of elements of input matrix: %d\n", A.m * A.n);
        printf("Number of elements of output matrix: %d\n", m * n);
        exit(1);
    }
    mtrx B = initm(m, n);
    for (int q = 0; q < m * n; q++) B.M[q / n][q % n] = A.M[q / A.n][q % A.n];
    return B;
}

vec gaussian(mtrx A, vec b)
{
    /* Gaussian elimination with partial pivoting and back substitution on [A | b]. */
    const int n = b.n;
    double **a = allocm(n, n + 1);
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) a[i][j] = A.M[i][j];
        a[i][n] = b.v[i];
    }
    for (int i = 0; i < n - 1; i++) {
        for (int k = i + 1; k < n; k++)
            if (fabs(a[i][i]) < fabs(a[k][i])) {
                double *t = a[i];
                a[i] = a[k];
                a[k] = t;
            }
        for (int k = i + 1; k < n; k++) {
            const double term = a[k][i] / a[i][i];
            for (int j = 0; j <= n; j++) a[k][j] = a[k][j] - term * a[i][j];
        }
    }
    vec x;
    x.n = n;
    x.v = allocv(n);
    for (int i = n - 1; i >= 0; i--) {
        double s = a[i][n];
        for (int j = i + 1; j < n; j++) s = s - a[i][j] * x.v[j];
        x.v[i] = s / a[i][i];
    }
    mtrx tmp = {a, n, n + 1};
    freem(tmp);
    return x;
}

void zerosv(vec v)
{
    for (int i = 0; i < v.n; i++) v.v[i] = 0;
}

double *allocv(int n)
{
    if (n < 1) fatal("** Error: invalid parameter **\n");
    double *v = (double *)malloc((size_t)n * sizeof(double));
    if (!v) fatal("** Error: insufficient memory **");
    return v;
}

double *freev(vec v)
{
    free(v.v);
    return NULL;
}

double **readm(char *filename, int *m, int *n)
{
    FILE *f = fopen(filename, "r");
    if (!f || fscanf(f, "%d %d", m, n) != 2) {
        printf("** Error: cannot read matrix file %s **\n", filename);
        exit(1);
    }
    double **A = allocm(*m, *n);
    for (int i = 0; i < *m; i++)
        for (int j = 0; j < *n; j++)
            if (fscanf(f, "%lf", &A[i][j]) != 1) A[i][j] = 0;
    fclose(f);
    return A;
}

double *readv(char *filename, int *n)
{
    FILE *f = fopen(filename, "r");
    if (!f || fscanf(f, "%d", n) != 1) {
        printf("** Error: cannot read vector file %s **\n", filename);
        exit(1);
    }
    double *v = allocv(*n);
    for (int i = 0; i < *n; i++)
        if (fscanf(f, "%lf", &v[i]) != 1) v[i] = 0;
    fclose(f);
    return v;
}

void printm(mtrx A)
{
    for (int i = 0; i < A.m; i++) {
        for (int j = 0; j < A.n; j++) printf(" %.4lf ", A.M[i][j]);
        printf("\n");
    }
}

void printv(vec v)
{
    printf("[");
    for (int i = 0; i < v.n; i++) printf(" %.4lf ", v.v[i]);
    printf(" ]\n");
}
