/*
 * fdtables.c -- finite-difference operator tables (host).
 *
 * The coefficients are those of the reference's Diff1/Diff2 (src/finitediff.c:51-292), and each
 * is computed with the same expression: the reference writes e.g. "(double)-1 / 60 / dx",
 * "-0.5 / dx" or "2. / (dx * dx)", i.e. ((num / den) / hp) with hp = h or h*h; the table below
 * stores (num, den) pairs and evaluates exactly that, so the doubles are identical.
 *   Diff1: one-sided [-1, 1]/h at the walls, then the 2nd-order central row, then (order 6)
 *          the 4th-order row, then the interior stencil of the order; right-wall rows carry the
 *          left-wall coefficient sequences translated to end at column n-1.
 *   Diff2: [2, -5, 4, -1]/h^2 at the walls, then lower-order central rows; right-wall rows are
 *          the left-wall rows reflected, D[n-1-r][n-1-c] = D[r][c].
 */
#include <string.h>

#include "cnv_host.h"

typedef struct {
    int width;
    double num[7], den[7];
} fd_row;

/* rows 0..2 from the left wall, then the interior stencil, per operator and order */
static const fd_row D1_WALL = {2, {-1, 1}, {1, 1}};
static const fd_row D1_C2 = {3, {-0.5, 0, 0.5}, {1, 1, 1}};
static const fd_row D1_C4 = {5, {1, -2, 0, 2, -1}, {12, 3, 1, 3, 12}};
static const fd_row D1_C6 = {7, {-1, 3, -3, 0, 3, -3, 1}, {60, 20, 4, 1, 4, 20, 60}};
static const fd_row D2_WALL = {4, {2, -5, 4, -1}, {1, 1, 1, 1}};
static const fd_row D2_C2 = {3, {1, -2, 1}, {1, 1, 1}};
static const fd_row D2_C4 = {5, {-1, 4, -5, 4, -1}, {12, 3, 2, 3, 12}};
static const fd_row D2_C6 = {7, {1, -3, 3, -49, 3, -3, 1}, {90, 20, 2, 18, 2, 20, 90}};

int cnv_fd_dense(int which, int n, int order, double h, double *D)
{
    if (order != 2 && order != 4 && order != 6) return -1;
    if (n < order + 2) return -2;
    memset(D, 0, sizeof(double) * (size_t)n * (size_t)n);
    const double hp = (which == 1) ? h : h * h;
    const int nwall = order / 2;  /* wall rows with their own (lower-order) stencils */
    const fd_row *wall[3], *inner;
    if (which == 1) {
        wall[0] = &D1_WALL; wall[1] = &D1_C2; wall[2] = &D1_C4;
        inner = (order == 2) ? &D1_C2 : (order == 4) ? &D1_C4 : &D1_C6;
    } else {
        wall[0] = &D2_WALL; wall[1] = &D2_C2; wall[2] = &D2_C4;
        inner = (order == 2) ? &D2_C2 : (order == 4) ? &D2_C4 : &D2_C6;
    }
    for (int r = 0; r < nwall; r++)
        for (int m = 0; m < wall[r]->width; m++)
            D[(size_t)r * n + m] = (wall[r]->num[m] / wall[r]->den[m]) / hp;
    const int half = inner->width / 2;
    for (int r = nwall; r < n - nwall; r++)
        for (int m = 0; m < inner->width; m++)
            D[(size_t)r * n + (r - half + m)] = (inner->num[m] / inner->den[m]) / hp;
    for (int r = 0; r < nwall; r++) {
        const int w = wall[r]->width;
        for (int m = 0; m < w; m++) {
            if (which == 1) /* translated copy ending at column n-1 */
                D[(size_t)(n - 1 - r) * n + (n - w + m)] = D[(size_t)r * n + m];
            else            /* reflection */
                D[(size_t)(n - 1 - r) * n + (n - 1 - m)] = D[(size_t)r * n + m];
        }
    }
    return 0;
}

int cnv_fd_windows(int n, const double *D, int *start, double *coef)
{
    if (n < 7) return -1;
    for (int r = 0; r < n; r++) {
        int s = r - 3;
        if (s < 0) s = 0;
        if (s > n - 7) s = n - 7;
        start[r] = s;
        for (int c = 0; c < n; c++) {
            const double a = D[(size_t)r * n + c];
            if (c >= s && c < s + 7)
                coef[(size_t)r * 7 + (c - s)] = a;
            else if (a != 0.0)
                return -1;
        }
    }
    return 0;
}
