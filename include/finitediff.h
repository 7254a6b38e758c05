/* finitediff.h -- finite-difference operator matrices, same C API as the reference
 * (reference include/finitediff.h:8-9).  Dense n x n matrices with the reference's coefficients;
 * an order other than 2, 4 or 6 prints the reference's message and exits (src/finitediff.c:150). */
#ifndef FINITEDIFF_H_INCLUDED
#define FINITEDIFF_H_INCLUDED

#include "linearalg.h"

mtrx Diff1(int n, int o, double dx); /* first derivative, order o (src/finitediff.c:51) */
mtrx Diff2(int n, int o, double dx); /* second derivative, order o (src/finitediff.c:178) */

#endif
