/*
 * qbf_quadrature.c -- quadrature tables of the collision integrals, runtime grid size N.
 *
 * Host-side setup, evaluated once per run and uploaded to the GPU by the engine.
 * The formulas are fixed by the numerical scheme (Lu & Mendl, JCP 291 (2015)) and
 * follow the reference's definitions so that results agree with it:
 *   I1: src/quadrature.c:150-238   psiR1 = R sin^2(x)/x on the 2N grid, angle theta1_i = i pi/J;
 *                                  psiR2 = R sin^2(x)/x on the N grid, theta2_j = theta1_j + pi/2J,
 *                                  times the principal-value weight -(2pi/J)^2 / cos(theta1_i - theta2_j);
 *                                  term index i + J*j (psiR1 repeated over j, psiR2 over i).
 *   I2: src/quadrature.c:245-286   2R sinc(s k.e_theta), (2R pi/J) sinc(s k.e_theta_perp), s = R pi/L.
 *   I3: src/quadrature.c:395-466   exp(i rho_m pi/L k.e_theta) on the N grid,
 *                                  (2R w_m pi/J) sinc(pi R/L k.e_theta_perp) on the 2N grid, index m + j*M.
 *   I4: src/quadrature.c:473-527   as I2 but the first table on the 2N grid.
 *   Gauss-Legendre: src/quadrature.c:293-388 (Newton on the three-term recurrence,
 *                                  nodes in descending order, same stopping rule).
 * Frequency axes: N grid [0..N/2-1, -N/2..-1], 2N grid [0..N-1, -N..-1]; the slow
 * index l multiplies cos(theta) (MATLAB meshgrid convention of the reference).
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

static double sinc_r(double x) { return fabs(x) < 1e-8 ? 1.0 : sin(x) / x; }
static double sin2_over_x(double x) { return fabs(x) < 1e-8 ? x : (sin(x) * sin(x)) / x; }
static int freq_n(int k, int n) { return k < n / 2 ? k : k - n; }

void qbf_legendre_gauss(unsigned M, double a, double b, double *x, double *w)
{
	double *y = malloc(sizeof(double) * M), *y0 = malloc(sizeof(double) * M);
	double *Lp = malloc(sizeof(double) * M);
	double *Lt = malloc(sizeof(double) * (size_t)M * (M + 1));   /* Lt[i + k*M] = P_k(y_i) */
	for (unsigned i = 0; i < M; i++) {
		const double xu = -1.0 + 2.0 * i / (double)(M - 1);
		y[i] = cos((2 * i + 1) * M_PI / (2 * M)) + 0.27 / M * sin(M_PI * xu * (M - 1) / (M + 1));
		y0[i] = 2;
	}
	double delta = 0;
	for (unsigned i = 0; i < M; i++) delta = fmax(delta, fabs(y[i] - y0[i]));
	while (delta > DBL_EPSILON) {
		for (unsigned i = 0; i < M; i++) {
			Lt[i] = 1;
			Lt[i + M] = y[i];
		}
		for (unsigned k = 1; k < M; k++)
			for (unsigned i = 0; i < M; i++)
				Lt[i + (k + 1) * M] = ((2 * k + 1) * y[i] * Lt[i + k * M] - k * Lt[i + (k - 1) * M]) / (k + 1);
		for (unsigned i = 0; i < M; i++)
			Lp[i] = (M + 1) * (Lt[i + (M - 1) * M] - y[i] * Lt[i + M * M]) / (1 - y[i] * y[i]);
		memcpy(y0, y, sizeof(double) * M);
		delta = 0;
		for (unsigned i = 0; i < M; i++) {
			const double step = Lt[i + M * M] / Lp[i];
			y[i] -= step;
			delta = fmax(delta, fabs(step));
		}
	}
	for (unsigned i = 0; i < M; i++) {
		x[i] = (a * (1 - y[i]) + b * (1 + y[i])) / 2;
		const double r = (double)(M + 1) / M;
		w[i] = (b - a) / ((1 - y[i] * y[i]) * (Lp[i] * Lp[i])) * (r * r);
	}
	free(Lt); free(Lp); free(y0); free(y);
}

void qbf_fourier_I1(int N, unsigned J, double L, double R, double *psiR1, double *psiR2)
{
	const int P = 2 * N;
	const size_t nn = (size_t)N * N, pp = (size_t)P * P;
	const double dth = M_PI / J, s = R * M_PI / (2 * L);
	for (unsigned i = 0; i < J; i++) {
		const double th = i * dth, c = cos(th), sn = sin(th);
		double *g = psiR1 + (size_t)i * pp;
		for (int l = 0; l < P; l++)
			for (int k = 0; k < P; k++)
				g[k + (size_t)P * l] = R * sin2_over_x(s * (freq_n(l, P) * c + freq_n(k, P) * sn));
		for (unsigned j = 1; j < J; j++) memcpy(psiR1 + (size_t)(i + J * j) * pp, g, sizeof(double) * pp);
	}
	for (unsigned j = 0; j < J; j++) {
		const double th2 = j * dth + 0.5 * dth, c = cos(th2), sn = sin(th2);
		double base[4096 * 4];
		double *b = nn <= sizeof(base) / sizeof(base[0]) ? base : malloc(sizeof(double) * nn);
		for (int l = 0; l < N; l++)
			for (int k = 0; k < N; k++)
				b[k + (size_t)N * l] = R * sin2_over_x(s * (freq_n(l, N) * c + freq_n(k, N) * sn));
		for (unsigned i = 0; i < J; i++) {
			const double wgt = -((2 * M_PI / J) * (2 * M_PI / J)) / cos(i * dth - th2);
			double *g = psiR2 + (size_t)(i + J * j) * nn;
			for (size_t e = 0; e < nn; e++) g[e] = b[e] * wgt;
		}
		if (b != base) free(b);
	}
}

void qbf_fourier_I2(int N, unsigned J, double L, double R, double *psiR1, double *psiR2)
{
	const size_t nn = (size_t)N * N;
	const double s = R * M_PI / L, f1 = 2 * R, f2 = 2 * R * M_PI / J;
	for (unsigned j = 0; j < J; j++) {
		const double th = j * (M_PI / J), c = cos(th), sn = sin(th);
		for (int l = 0; l < N; l++)
			for (int k = 0; k < N; k++) {
				const int kl = freq_n(l, N), kk = freq_n(k, N);
				psiR1[(size_t)j * nn + k + (size_t)N * l] = f1 * sinc_r(s * (kl * c + kk * sn));
				psiR2[(size_t)j * nn + k + (size_t)N * l] = f2 * sinc_r(s * (-kl * sn + kk * c));
			}
	}
}

void qbf_fourier_I3(int N, unsigned J, double L, double R, unsigned M, double *psiR1, double *psiR2)
{
	const int P = 2 * N;
	const size_t nn = (size_t)N * N, pp = (size_t)P * P;
	double *rho = malloc(sizeof(double) * M), *wm = malloc(sizeof(double) * M);
	qbf_legendre_gauss(M, -R, R, rho, wm);
	const double s2 = M_PI * R / L;
	for (unsigned j = 0; j < J; j++) {
		const double th = j * (M_PI / J), c = cos(th), sn = sin(th);
		for (unsigned m = 0; m < M; m++) {
			const size_t q = m + (size_t)j * M;
			const double s1 = rho[m] * M_PI / L, f2 = 2 * R * wm[m] * M_PI / J;
			for (int l = 0; l < N; l++)
				for (int k = 0; k < N; k++) {
					const double t = s1 * (freq_n(l, N) * c + freq_n(k, N) * sn);
					psiR1[2 * (q * nn + k + (size_t)N * l)] = cos(t);
					psiR1[2 * (q * nn + k + (size_t)N * l) + 1] = sin(t);
				}
			for (int l = 0; l < P; l++)
				for (int k = 0; k < P; k++)
					psiR2[q * pp + k + (size_t)P * l] = f2 * sinc_r(s2 * (-freq_n(l, P) * sn + freq_n(k, P) * c));
		}
	}
	free(wm); free(rho);
}

void qbf_fourier_I4(int N, unsigned J, double L, double R, double *psiR1, double *psiR2)
{
	const int P = 2 * N;
	const size_t nn = (size_t)N * N, pp = (size_t)P * P;
	const double s = M_PI * R / L, f1 = 2 * R, f2 = 2 * R * M_PI / J;
	for (unsigned j = 0; j < J; j++) {
		const double th = j * (M_PI / J), c = cos(th), sn = sin(th);
		for (int l = 0; l < P; l++)
			for (int k = 0; k < P; k++)
				psiR1[(size_t)j * pp + k + (size_t)P * l] = f1 * sinc_r(s * (freq_n(l, P) * c + freq_n(k, P) * sn));
		for (int l = 0; l < N; l++)
			for (int k = 0; k < N; k++)
				psiR2[(size_t)j * nn + k + (size_t)N * l] = f2 * sinc_r(s * (-freq_n(l, N) * sn + freq_n(k, N) * c));
	}
}
