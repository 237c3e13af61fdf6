/*
 * quadrature.h -- quadrature tables of the collision integrals (drop-in for the
 * reference's include/quadrature.h:35-95; same struct layout, callers fill psiR1 /
 * psiR2 directly, e.g. test/I2_test.c:80-93).
 */
#ifndef QUADRATURE_H
#define QUADRATURE_H

#include "grid.h"

typedef struct { unsigned int num; grid2_t *psiR1; grid_t *psiR2; } quadI1_t;
void QuadI1_Allocate(const unsigned int num, quadI1_t *quad);
void QuadI1_Delete(quadI1_t *quad);

typedef struct { unsigned int num; grid_t *psiR1; grid_t *psiR2; } quadI2_t;
void QuadI2_Allocate(const unsigned int J, quadI2_t *quad);
void QuadI2_Delete(quadI2_t *quad);

typedef struct { unsigned int num; cgrid_t *psiR1; grid2_t *psiR2; } quadI3_t;
void QuadI3_Allocate(const unsigned int num, quadI3_t *quad);
void QuadI3_Delete(quadI3_t *quad);

void FourierI1(const unsigned int J, const double L, const double R, quadI1_t *quad);
void FourierI2(const unsigned int J, const double L, const double R, quadI2_t *quad);
void FourierI3(const unsigned int J, const double L, const double R, const unsigned int M, quadI3_t *quad);
void FourierI4(const unsigned int J, const double L, const double R, quadI1_t *quad);

/* Gauss-Legendre nodes (descending) and weights on [a, b] (src/quadrature.c:293) */
void LegendreGaussQuad(const unsigned int M, const double a, const double b, double *x, double *w);

#endif
