/* linearalg.h -- the reference's matrix container and dense helpers, same C API
 * (reference include/linearalg.h:6-43).  Row-pointer storage, one allocation per row, so that
 * matrices returned by this library are freed by freem() exactly like the reference's. */
#ifndef LINEARALG_H_INCLUDED
#define LINEARALG_H_INCLUDED

typedef struct matrix {
    double **M; /* M[i][j], row i */
    int m;      /* rows */
    int n;      /* columns */
} mtrx;

typedef struct vector {
    double *v;
    int n;
} vec;

void zerosm(mtrx A);
double **allocm(int m, int n);
double **freem(mtrx A);                         /* returns NULL: x.M = freem(x) */
double **readm(char *filename, int *m, int *n);
void printm(mtrx A);
void zerosv(vec v);
double *allocv(int n);
double *freev(vec v);
double *readv(char *filename, int *n);
void printv(vec v);
mtrx mtrxmul(mtrx A, mtrx B);                   /* C = A B, k summed in ascending order from 0.0 */
vec gaussian(mtrx A, vec b);
mtrx kronecker(mtrx A, mtrx B);
mtrx reshape(mtrx A, int m, int n);             /* row-major reflow */
mtrx eye(int n);
mtrx initm(int m, int n);                       /* zero-initialised m x n */
void invsig(mtrx A);                            /* A = -A */
double maxel(mtrx A);
double minel(mtrx A);
void mtrxcpy(mtrx A, mtrx B);                   /* A = B */
void set_openmp_config(int enabled);

#endif
