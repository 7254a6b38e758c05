/* cnv_host.h -- internal host helpers shared by the C legacy layer and the HIP context. */
#ifndef CNV_HOST_H
#define CNV_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

/* Dense n x n first (which = 1) or second (which = 2) derivative operator with the reference's
 * coefficients (src/finitediff.c:51-292) into D (row-major, zero-filled here).
 * Returns 0, or -1 for an order other than 2/4/6, or -2 when n is too small for the order. */
int cnv_fd_dense(int which, int n, int order, double h, double *D);

/* Row windows of a dense FD operator: start[r] = clamp(r-3, 0, n-7), coef[r*7+m] = D[r][start+m].
 * Returns 0, or -1 when a nonzero lies outside its row's window (cannot happen for n >= 8). */
int cnv_fd_windows(int n, const double *D, int *start, double *coef);

#ifdef __cplusplus
}
#endif
#endif
