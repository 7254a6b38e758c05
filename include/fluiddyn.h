/* fluiddyn.h -- time integration and diagnostics, same C API as the reference
 * (reference include/fluiddyn.h:8-14), evaluated on the GPU. */
#ifndef FLUIDDYN_H_INCLUDED
#define FLUIDDYN_H_INCLUDED

#include "linearalg.h"

/* w = (-u*dwdx - v*dwdy + (1./Re)*(d2wdx2 + d2wdy2))*dt + w at every point (src/fluiddyn.c:71,98) */
void euler(mtrx w, mtrx dwdx, mtrx dwdy, mtrx d2wdx2, mtrx d2wdy2, mtrx u, mtrx v, double Re, double dt);
mtrx continuity(mtrx dudx, mtrx dvdy);      /* dudx + dvdy (src/fluiddyn.c:126) */
mtrx vorticity(mtrx dvdx, mtrx dudy);       /* the 2nd argument minus the 1st (src/fluiddyn.c:179) */
mtrx pressure(void);                        /* declared only, as in the reference (include/fluiddyn.h:11) */
void set_fluiddyn_openmp_config(int enabled); /* accepted; the GPU path is used (src/fluiddyn.c:35) */

#endif
