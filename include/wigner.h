/*
 * wigner.h -- Wigner state in Pauli representation (drop-in for the reference's
 * include/wigner.h:35-50).  Velocity grid [0:N/2-1, -N/2:-1] * 2L/N per direction.
 */
#ifndef WIGNER_H
#define WIGNER_H

#include "grid.h"

typedef struct { grid_t comp[4]; } wignerV_t;    /* physical velocity space */
typedef struct { cgrid_t comp[4]; } wignerF_t;   /* Fourier space */

#endif
