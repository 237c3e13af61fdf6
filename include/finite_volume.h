/*
 * finite_volume.h -- slope-limited finite-volume transport steps (drop-in for the
 * reference's include/finite_volume.h:29-40).
 */
#ifndef FINITE_VOLUME_H
#define FINITE_VOLUME_H

#ifndef restrict_
#ifdef __cplusplus
#define restrict_ __restrict__
#else
#define restrict_ restrict
#endif
#endif

void SlopeLimiterStepPeriodic(const double h, const double dt, const double A, const unsigned int N, const double *restrict_ Un, double *restrict_ Un1);
void SlopeLimiterStepDirichlet(const double h, const double dt, const double A, const unsigned int N, const double *restrict_ Un, double *restrict_ Un1);
double CalculateRightwardFlux(const unsigned int m, const double *restrict_ A, const double *restrict_ U);
double CalculateLeftwardFlux (const unsigned int m, const double *restrict_ A, const double *restrict_ U);
void SlopeLimiterStepMaxwell(const double h, const double dt, const unsigned int m, const double *restrict_ A, const unsigned int N,
	const double lambda, const double fluxL, const double fluxR, const double *restrict_ UmaxwL, const double *restrict_ UmaxwR,
	const double *restrict_ Un, double *restrict_ Un1);

#endif
