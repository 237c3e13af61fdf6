/*
 * collision.h -- conservative (Cc) and dissipative (Cd) collision operators
 * (drop-in for the reference's include/collision.h:47-73).
 */
#ifndef COLLISION_H
#define COLLISION_H

#include "integrals.h"
#include "wigner.h"

typedef struct { void *impl; } CdInterm_t;
void CdInterm_Create(CdInterm_t *interm);
void CdInterm_Delete(CdInterm_t *interm);
void CdInt(const wignerF_t *restrict_ W, const quadI2_t *restrict_ quadI2, const quadI3_t *restrict_ quadI3,
	const quadI1_t *restrict_ quadI4, CdInterm_t *restrict_ interm, wignerF_t *restrict_ Cd);

typedef struct { void *impl; } CcInterm_t;
void CcInterm_Create(CcInterm_t *interm);
void CcInterm_Delete(CcInterm_t *interm);
void CcInt(const wignerF_t *restrict_ W, const quadI1_t *restrict_ quadI1, CcInterm_t *restrict_ interm,
	wignerF_t *restrict_ Cc);

#endif
