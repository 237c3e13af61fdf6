/*
 * integrals.h -- the I1, I2, I3 collision integrals (drop-in for the reference's
 * include/integrals.h:56-138).  The interm structs are private to the library:
 * callers only allocate them (usually on the stack) and pass them back.
 */
#ifndef INTEGRALS_H
#define INTEGRALS_H

#include "grid.h"
#include "quadrature.h"

typedef struct { void *impl; } I1interm_t;
typedef struct { void *impl; } I2interm_t;
typedef struct { void *impl; } I3interm_t;

void I1interm_Create(I1interm_t *interm);
void I1interm_Delete(I1interm_t *interm);
void I1integral(const cgrid_t *restrict_ f, const cgrid_t *restrict_ g, const cgrid_t *restrict_ h,
	const quadI1_t *restrict_ quad, I1interm_t *restrict_ interm, cgrid_t *restrict_ I1);

void I2interm_Create(I2interm_t *interm);
void I2interm_Delete(I2interm_t *interm);
void I2integral(const cgrid_t *restrict_ f, const cgrid_t *restrict_ g, const cgrid_t *restrict_ h,
	const quadI2_t *restrict_ quad, I2interm_t *restrict_ interm, cgrid_t *restrict_ I2);

void I3interm_Create(I3interm_t *interm);
void I3interm_Delete(I3interm_t *interm);
void I3integral(const cgrid_t *restrict_ f, const cgrid_t *restrict_ g, const cgrid_t *restrict_ h,
	const quadI3_t *restrict_ quad, I3interm_t *restrict_ interm, cgrid_t *restrict_ I3);

#endif
