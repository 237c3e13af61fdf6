/*
 * orc_transport.c -- slope-limited finite-volume transport for the oracle
 * (TEST INFRASTRUCTURE ONLY).
 *
 * Restates src/finite_volume.c: minmod (:35-38, LeVeque (16.52)), the periodic
 * step (:52-104, LeVeque (16.57)), the Dirichlet step (:118-179), the flux
 * helpers (:186-216) and the Maxwell diffuse-reflection step (:241-323).
 *
 * Parity: pinned by test/data/Un1_periodic_ref.dat, Un1_Dirichlet_ref.dat
 * (h=0.08, dt=0.02, A=-1.2, N=128) and Un1_Maxwell_ref.dat
 * (test/slope_lim_test2.c parameters); invariants: the periodic step
 * conserves sum(U) and maps constants to themselves.
 */
#include "qbf_oracle.h"
#include <math.h>
#include <stdlib.h>

static inline double minmod(double a, double b)
{
	return 0.5 * (copysign(1, a) + copysign(1, b)) * fmin(fabs(a), fabs(b));
}

void orc_SlopeLimiterStepPeriodic(double h, double dt, double A, unsigned N, const double *Un, double *Un1)
{
	const double Aabs = fabs(A);
	const double nu = dt * A / h;
	const double S = 0.5 * A * (copysign(1, nu) - nu);
	const double Sabs = 0.5 * Aabs * (copysign(1, nu) - nu);
	double *alpha = malloc(sizeof(double) * N), *hb = malloc(sizeof(double) * N);
	for (unsigned j = 0; j < N - 1; j++) alpha[j] = Un[j + 1] - Un[j];
	alpha[N - 1] = Un[0] - Un[N - 1];
	for (unsigned j = 1; j < N; j++) hb[j] = minmod(alpha[j], alpha[j - 1]);
	hb[0] = minmod(alpha[0], alpha[N - 1]);
	const double scale = dt / (2 * h);
	for (unsigned j = 0; j < N; j++) {
		const unsigned jp = (j + 1) % N, jm = (j + N - 1) % N;
		Un1[j] = Un[j] + scale * (
			-A * (Un[jp] - Un[jm]) + Aabs * (Un[jp] - 2 * Un[j] + Un[jm])
			- S * (hb[jp] - hb[jm]) + Sabs * (hb[jp] - 2 * hb[j] + hb[jm]));
	}
	free(hb); free(alpha);
}

void orc_SlopeLimiterStepDirichlet(double h, double dt, double A, unsigned N, const double *Un, double *Un1)
{
	const double Aabs = fabs(A);
	const double nu = dt * A / h;
	const double S = 0.5 * A * (copysign(1, nu) - nu);
	const double Sabs = 0.5 * Aabs * (copysign(1, nu) - nu);
	double *alpha = malloc(sizeof(double) * N), *hb = malloc(sizeof(double) * N);
	for (unsigned j = 0; j < N - 1; j++) alpha[j] = Un[j + 1] - Un[j];
	alpha[N - 1] = 0;
	for (unsigned j = 1; j < N - 1; j++) hb[j] = minmod(alpha[j], alpha[j - 1]);
	hb[0] = 0;
	hb[N - 1] = 0;
	const double scale = dt / (2 * h);
	for (unsigned j = 1; j < N - 1; j++)
		Un1[j] = Un[j] + scale * (
			-A * (Un[j + 1] - Un[j - 1]) + Aabs * (Un[j + 1] - 2 * Un[j] + Un[j - 1])
			- S * (hb[j + 1] - hb[j - 1]) + Sabs * (hb[j + 1] - 2 * hb[j] + hb[j - 1]));
	if (A >= 0) {
		Un1[0] = Un[0];
		Un1[N - 1] = Un[N - 1] + 2 * scale * (-A * (Un[N - 1] - Un[N - 2]) - S * (hb[N - 1] - hb[N - 2]));
	} else {
		Un1[0] = Un[0] + 2 * scale * (-A * (Un[1] - Un[0]) - S * (hb[1] - hb[0]));
		Un1[N - 1] = Un[N - 1];
	}
	free(hb); free(alpha);
}

double orc_CalculateRightwardFlux(unsigned m, const double *A, const double *U)
{
	double flux = 0;
	for (unsigned i = 0; i < m / 2; i++) flux += A[i] * U[i];
	return flux;
}

double orc_CalculateLeftwardFlux(unsigned m, const double *A, const double *U)
{
	double flux = 0;
	for (unsigned i = m / 2; i < m; i++) flux += fabs(A[i]) * U[i];
	return flux;
}

void orc_SlopeLimiterStepMaxwell(double h, double dt, unsigned m, const double *A, unsigned N,
	double lambda, double fluxL, double fluxR, const double *UmaxwL, const double *UmaxwR,
	const double *Un, double *Un1)
{
	double *Um1 = malloc(sizeof(double) * m), *UN1 = malloc(sizeof(double) * m);
	for (unsigned i = 0; i < m / 2; i++) {
		UN1[i] = Un[i + m * (N - 1)];
		Um1[i] = (1 - lambda) * Un[m - 1 - i] + lambda * fluxL * UmaxwL[i];
	}
	for (unsigned i = m / 2; i < m; i++) {
		Um1[i] = Un[i];
		UN1[i] = (1 - lambda) * Un[m - 1 - i + m * (N - 1)] + lambda * fluxR * UmaxwR[i];
	}
	/* V[j] for j = -2..N+1 stored at V[j+2]; alpha[j] j=-2..N at alpha[j+2]; hb[j] j=-1..N at hb[j+1] */
	double *V = malloc(sizeof(double) * (N + 4));
	double *alpha = malloc(sizeof(double) * (N + 3));
	double *hb = malloc(sizeof(double) * (N + 2));
	for (unsigned i = 0; i < m; i++) {
		const double Aabs = fabs(A[i]);
		const double nu = dt * A[i] / h;
		const double S = 0.5 * A[i] * (copysign(1, nu) - nu);
		const double Sabs = 0.5 * Aabs * (copysign(1, nu) - nu);
		V[0] = Um1[i];
		V[1] = Um1[i];
		for (unsigned j = 0; j < N; j++) V[j + 2] = Un[i + m * j];
		V[N + 2] = UN1[i];
		V[N + 3] = UN1[i];
		for (int j = -2; j <= (int)N; j++) alpha[j + 2] = V[j + 3] - V[j + 2];
		for (int j = -1; j <= (int)N; j++) hb[j + 1] = minmod(alpha[j + 2], alpha[j + 1]);
		const double scale = dt / (2 * h);
		for (int j = 0; j < (int)N; j++) {
			const double vp = V[j + 3], v0 = V[j + 2], vm = V[j + 1];
			const double bp = hb[j + 2], b0 = hb[j + 1], bm = hb[j];
			Un1[i + m * j] = v0 + scale * (
				-A[i] * (vp - vm) + Aabs * (vp - 2 * v0 + vm)
				- S * (bp - bm) + Sabs * (bp - 2 * b0 + bm));
		}
	}
	free(hb); free(alpha); free(V); free(UN1); free(Um1);
}
