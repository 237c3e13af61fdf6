/*
 * grid.h -- square grids with real or complex entries (drop-in for the reference's
 * include/grid.h:34-102).  N_GRID keeps the reference's compile-time default of 32;
 * build with -DN_GRID=64 against a library built for the same N.
 */
#ifndef GRID_H
#define GRID_H

#include <fftw3.h>

#ifndef restrict_
#ifdef __cplusplus
#define restrict_ __restrict__
#else
#define restrict_ restrict
#endif
#endif

#ifndef N_GRID
#define N_GRID 32                 /* number of grid points in each direction (power of two) */
#endif
#define N_MASK (N_GRID - 1)

typedef struct { double data[N_GRID * N_GRID]; } grid_t;             /* real N x N */
typedef struct { double data[4 * N_GRID * N_GRID]; } grid2_t;        /* real 2N x 2N */
typedef struct { fftw_complex data[N_GRID * N_GRID]; } cgrid_t;      /* complex N x N */
typedef struct { fftw_complex data[4 * N_GRID * N_GRID]; } cgrid2_t; /* complex 2N x 2N */

void MultiplyPointwiseRR (const   grid_t *restrict_ f, const   grid_t *restrict_ g,   grid_t *restrict_ ret);
void MultiplyPointwiseRC (const   grid_t *restrict_ f, const  cgrid_t *restrict_ g,  cgrid_t *restrict_ ret);
void MultiplyPointwiseCC (const  cgrid_t *restrict_ f, const  cgrid_t *restrict_ g,  cgrid_t *restrict_ ret);
void MultiplyPointwiseRR2(const  grid2_t *restrict_ f, const  grid2_t *restrict_ g,  grid2_t *restrict_ ret);
void MultiplyPointwiseCC2(const cgrid2_t *restrict_ f, const cgrid2_t *restrict_ g, cgrid2_t *restrict_ ret);

void MultiplyPointwiseAssignRR (const   grid_t *restrict_ f,   grid_t *restrict_ ret);
void MultiplyPointwiseAssignRC (const   grid_t *restrict_ f,  cgrid_t *restrict_ ret);
void MultiplyPointwiseAssignRC2(const  grid2_t *restrict_ f, cgrid2_t *restrict_ ret);
void MultiplyPointwiseAssignCC2(const cgrid2_t *restrict_ f, cgrid2_t *restrict_ ret);

void GridZeroPad(const cgrid_t *restrict_ f, cgrid2_t *restrict_ ret);
void GridExtract(const cgrid2_t *restrict_ f, cgrid_t *restrict_ ret);

#endif
