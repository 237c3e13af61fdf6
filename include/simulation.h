/*
 * simulation.h -- collision step, inhomogeneous Strang step and simulation loops
 * (drop-in for the reference's include/simulation.h:46-129).
 */
#ifndef SIMULATION_H
#define SIMULATION_H

#include "collision.h"

typedef struct { void *impl; } collisionInterm_t;
void CollisionInterm_Create(collisionInterm_t *interm);
void CollisionInterm_Delete(collisionInterm_t *interm);
void CollisionStep(const wignerF_t *restrict_ W, const quadI1_t *restrict_ quadI1, const quadI2_t *restrict_ quadI2,
	const quadI3_t *restrict_ quadI3, const quadI1_t *restrict_ quadI4,
	const double dt, const double Bext[3], collisionInterm_t *restrict_ interm, wignerF_t *restrict_ Wnext);

typedef struct {
	unsigned int numVol;      /* number of finite volumes */
	double L;                 /* velocity box half-width */
	void *impl;
} inhomStepInterm_t;
void InhomStepInterm_Create(const unsigned int numVol, const double L, inhomStepInterm_t *interm);
void InhomStepInterm_Delete(inhomStepInterm_t *interm);

void InhomStepPeriodic(const wignerV_t *restrict_ W, const quadI1_t *restrict_ quadI1, const quadI2_t *restrict_ quadI2,
	const quadI3_t *restrict_ quadI3, const quadI1_t *restrict_ quadI4,
	const double h, const double dt, const double *restrict_ Bext, inhomStepInterm_t *restrict_ interm, wignerV_t *restrict_ Wnext);
void InhomStepDirichlet(const wignerV_t *restrict_ W, const quadI1_t *restrict_ quadI1, const quadI2_t *restrict_ quadI2,
	const quadI3_t *restrict_ quadI3, const quadI1_t *restrict_ quadI4,
	const double h, const double dt, const double *restrict_ Bext, inhomStepInterm_t *restrict_ interm, wignerV_t *restrict_ Wnext);
void InhomStepMaxwell(const wignerV_t *restrict_ W, const quadI1_t *restrict_ quadI1, const quadI2_t *restrict_ quadI2,
	const quadI3_t *restrict_ quadI3, const quadI1_t *restrict_ quadI4,
	const double h, const double dt, const double lambda, const wignerV_t *restrict_ WMaxwL, const wignerV_t *restrict_ WMaxwR,
	const double *restrict_ Bext, inhomStepInterm_t *restrict_ interm, wignerV_t *restrict_ Wnext);

typedef struct { void *impl; } simulationHomInterm_t;
void SimulationHomInterm_Create(simulationHomInterm_t *interm);
void SimulationHomInterm_Delete(simulationHomInterm_t *interm);
void SimulationHomogeneous(const wignerV_t *restrict_ W0, const quadI1_t *restrict_ quadI1, const quadI2_t *restrict_ quadI2,
	const quadI3_t *restrict_ quadI3, const quadI1_t *restrict_ quadI4,
	const double dt, const unsigned int numsteps, const double Bext[3], simulationHomInterm_t *restrict_ interm, wignerV_t *restrict_ Wevolv);

void SimulationInhomogeneousPeriodic(const wignerV_t *restrict_ W0, const quadI1_t *restrict_ quadI1, const quadI2_t *restrict_ quadI2,
	const quadI3_t *restrict_ quadI3, const quadI1_t *restrict_ quadI4,
	const double h, const double dt, const unsigned int numsteps, const double *restrict_ Bext, inhomStepInterm_t *restrict_ interm, wignerV_t *restrict_ Wevolv);
void SimulationInhomogeneousDirichlet(const wignerV_t *restrict_ W0, const quadI1_t *restrict_ quadI1, const quadI2_t *restrict_ quadI2,
	const quadI3_t *restrict_ quadI3, const quadI1_t *restrict_ quadI4,
	const double h, const double dt, const unsigned int numsteps, const double *restrict_ Bext, inhomStepInterm_t *restrict_ interm, wignerV_t *restrict_ Wevolv);
void SimulationInhomogeneousMaxwell(const wignerV_t *restrict_ W0, const quadI1_t *restrict_ quadI1, const quadI2_t *restrict_ quadI2,
	const quadI3_t *restrict_ quadI3, const quadI1_t *restrict_ quadI4,
	const double h, const double dt, const double lambda, const wignerV_t *restrict_ WMaxwL, const wignerV_t *restrict_ WMaxwR,
	const unsigned int numsteps, const double *restrict_ Bext, inhomStepInterm_t *restrict_ interm, wignerV_t *restrict_ Wevolv);

#endif
