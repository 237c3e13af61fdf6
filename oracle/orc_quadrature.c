/*
 * orc_quadrature.c -- quadrature tables for the oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Restates src/quadrature.c with the grid size N as an argument.  Formulas,
 * term ordering (index i + J*j for I1, m + j*M for I3) and the small-argument
 * branches of sinc / sin*sinc follow the reference line by line:
 *   sin_sinc, sinc ........ src/quadrature.c:116-142
 *   FourierI1 ............. src/quadrature.c:150-238
 *   FourierI2 ............. src/quadrature.c:245-286
 *   LegendreGaussQuad ..... src/quadrature.c:293-388 (Newton on the Legendre
 *                           recurrence, nodes in descending order)
 *   FourierI3 ............. src/quadrature.c:395-466
 *   FourierI4 ............. src/quadrature.c:473-527
 *
 * Parity: LegendreGaussQuad is pinned by test/data/x_ref.dat, w_ref.dat and by
 * numpy.polynomial.legendre.leggauss; the I1..I4 tables are pinned through the
 * integral fixtures I1_ref, I2_ref2, I3_ref, I4_ref, which were produced with
 * these tables (test/I*_test.c).
 */
#include "qbf_oracle.h"
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline double sq(double x) { return x * x; }

static inline double sin_sinc(double x)
{
	if (fabs(x) < 1e-8) return x;
	return sq(sin(x)) / x;
}

static inline double sinc(double x)
{
	if (fabs(x) < 1e-8) return 1;
	return sin(x) / x;
}

static void kgrids(int N, int *kg, int *kg2)
{
	for (int k = 0; k < N; k++) kg[k] = k < N / 2 ? k : k - N;
	for (int k = 0; k < 2 * N; k++) kg2[k] = k < N ? k : k - 2 * N;
}

void orc_FourierI1(int N, unsigned J, double L, double R, double *psiR1, double *psiR2)
{
	const int P = 2 * N;
	const size_t NN = (size_t)N * N, PP = (size_t)P * P;
	int *kg = malloc(sizeof(int) * N), *kg2 = malloc(sizeof(int) * P);
	kgrids(N, kg, kg2);
	double *th1 = malloc(sizeof(double) * J), *th2 = malloc(sizeof(double) * J);
	const double dth = M_PI / J;
	for (unsigned j = 0; j < J; j++) {
		th1[j] = j * dth;
		th2[j] = th1[j] + 0.5 * dth;
	}
	const double s = R * M_PI / (2 * L);
	for (unsigned i = 0; i < J; i++) {
		const double c = cos(th1[i]), sn = sin(th1[i]);
		double *dst = &psiR1[(size_t)i * PP];
		for (int l = 0; l < P; l++)
			for (int k = 0; k < P; k++)
				dst[k + (size_t)P * l] = R * sin_sinc(s * (kg2[l] * c + kg2[k] * sn));
		for (unsigned j = 1; j < J; j++)
			memcpy(&psiR1[(size_t)(i + J * j) * PP], dst, sizeof(double) * PP);
	}
	for (unsigned j = 0; j < J; j++) {
		const double c = cos(th2[j]), sn = sin(th2[j]);
		double *dst = &psiR2[(size_t)(J * j) * NN];
		for (int l = 0; l < N; l++)
			for (int k = 0; k < N; k++)
				dst[k + (size_t)N * l] = R * sin_sinc(s * (kg[l] * c + kg[k] * sn));
		for (unsigned i = 1; i < J; i++)
			memcpy(&psiR2[(size_t)(i + J * j) * NN], dst, sizeof(double) * NN);
	}
	for (unsigned j = 0; j < J; j++)
		for (unsigned i = 0; i < J; i++) {
			const double weight = -sq(2 * M_PI / J) / cos(th1[i] - th2[j]);
			double *dst = &psiR2[(size_t)(i + J * j) * NN];
			for (size_t e = 0; e < NN; e++) dst[e] *= weight;
		}
	free(th2); free(th1); free(kg2); free(kg);
}

void orc_FourierI2(int N, unsigned J, double L, double R, double *psiR1, double *psiR2)
{
	const size_t NN = (size_t)N * N;
	int *kg = malloc(sizeof(int) * N), *kg2 = malloc(sizeof(int) * 2 * N);
	kgrids(N, kg, kg2);
	const double dth = M_PI / J;
	const double f1 = 2 * R, f2 = 2 * R * M_PI / J, s = R * M_PI / L;
	for (unsigned j = 0; j < J; j++) {
		const double th = j * dth, c = cos(th), sn = sin(th);
		for (int l = 0; l < N; l++)
			for (int k = 0; k < N; k++) {
				psiR1[(size_t)j * NN + k + (size_t)N * l] = f1 * sinc(s * (kg[l] * c + kg[k] * sn));
				psiR2[(size_t)j * NN + k + (size_t)N * l] = f2 * sinc(s * (-kg[l] * sn + kg[k] * c));
			}
	}
	free(kg2); free(kg);
}

void orc_LegendreGaussQuad(unsigned M, double a, double b, double *x, double *w)
{
	const double invN = 1.0 / (M - 1);
	double *xu = malloc(sizeof(double) * M), *y = malloc(sizeof(double) * M);
	double *Lm = calloc((size_t)M * (M + 1), sizeof(double));
	double *Lp = calloc(M, sizeof(double));
	double *y0 = malloc(sizeof(double) * M);
	for (unsigned i = 0; i < M; i++) xu[i] = -1.0 + 2 * i * invN;
	for (unsigned i = 0; i < M; i++)
		y[i] = cos((2 * i + 1) * M_PI / (2 * M)) + 0.27 / M * sin(M_PI * xu[i] * (M - 1) / (M + 1));
	for (unsigned i = 0; i < M; i++) y0[i] = 2;
	double diff = 0;
	for (unsigned i = 0; i < M; i++) diff = fmax(diff, fabs(y[i] - y0[i]));
	while (diff > DBL_EPSILON) {
		for (unsigned i = 0; i < M; i++) { Lm[i] = 1; Lm[i + M] = y[i]; }
		for (unsigned k = 1; k < M; k++)
			for (unsigned i = 0; i < M; i++)
				Lm[i + (k + 1) * M] = ((2 * k + 1) * y[i] * Lm[i + k * M] - k * Lm[i + (k - 1) * M]) / (k + 1);
		for (unsigned i = 0; i < M; i++)
			Lp[i] = (M + 1) * (Lm[i + (M - 1) * M] - y[i] * Lm[i + M * M]) / (1 - sq(y[i]));
		memcpy(y0, y, sizeof(double) * M);
		diff = 0;
		for (unsigned i = 0; i < M; i++) {
			const double t = Lm[i + M * M] / Lp[i];
			y[i] -= t;
			diff = fmax(diff, fabs(t));
		}
	}
	for (unsigned i = 0; i < M; i++) {
		x[i] = (a * (1 - y[i]) + b * (1 + y[i])) / 2;
		w[i] = (b - a) / ((1 - sq(y[i])) * sq(Lp[i])) * sq((double)(M + 1) / M);
	}
	free(y0); free(Lp); free(Lm); free(y); free(xu);
}

void orc_FourierI3(int N, unsigned J, double L, double R, unsigned M, double *psiR1, double *psiR2)
{
	const int P = 2 * N;
	const size_t NN = (size_t)N * N, PP = (size_t)P * P;
	int *kg = malloc(sizeof(int) * N), *kg2 = malloc(sizeof(int) * P);
	kgrids(N, kg, kg2);
	double *rho = malloc(sizeof(double) * M), *w = malloc(sizeof(double) * M);
	orc_LegendreGaussQuad(M, -R, R, rho, w);
	const double dth = M_PI / J, s2 = M_PI * R / L;
	for (unsigned j = 0; j < J; j++) {
		const double th = j * dth, c = cos(th), sn = sin(th);
		for (unsigned m = 0; m < M; m++) {
			double *g1 = &psiR1[(size_t)(m + j * M) * 2 * NN];
			double *g2 = &psiR2[(size_t)(m + j * M) * PP];
			const double s1 = rho[m] * M_PI / L;
			const double f2 = 2 * R * w[m] * M_PI / J;
			for (int l = 0; l < N; l++)
				for (int k = 0; k < N; k++) {
					const double t = s1 * (kg[l] * c + kg[k] * sn);
					g1[2 * (k + (size_t)N * l)] = cos(t);
					g1[2 * (k + (size_t)N * l) + 1] = sin(t);
				}
			for (int l = 0; l < P; l++)
				for (int k = 0; k < P; k++)
					g2[k + (size_t)P * l] = f2 * sinc(s2 * (-kg2[l] * sn + kg2[k] * c));
		}
	}
	free(w); free(rho); free(kg2); free(kg);
}

void orc_FourierI4(int N, unsigned J, double L, double R, double *psiR1, double *psiR2)
{
	const int P = 2 * N;
	const size_t NN = (size_t)N * N, PP = (size_t)P * P;
	int *kg = malloc(sizeof(int) * N), *kg2 = malloc(sizeof(int) * P);
	kgrids(N, kg, kg2);
	const double dth = M_PI / J;
	const double f1 = 2 * R, f2 = 2 * R * M_PI / J, s = M_PI * R / L;
	for (unsigned j = 0; j < J; j++) {
		const double th = j * dth, c = cos(th), sn = sin(th);
		for (int l = 0; l < P; l++)
			for (int k = 0; k < P; k++)
				psiR1[(size_t)j * PP + k + (size_t)P * l] = f1 * sinc(s * (kg2[l] * c + kg2[k] * sn));
		for (int l = 0; l < N; l++)
			for (int k = 0; k < N; k++)
				psiR2[(size_t)j * NN + k + (size_t)N * l] = f2 * sinc(s * (-kg[l] * sn + kg[k] * c));
	}
	free(kg2); free(kg);
}
