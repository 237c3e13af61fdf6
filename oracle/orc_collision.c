/*
 * orc_collision.c -- integrals, collision operators and time stepping for the
 * oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Restates, with runtime N and the oracle's own FFT:
 *   GridZeroPad / GridExtract ......... src/grid.c:171-210
 *   I1integral (also I4) .............. src/integrals.c:90-133  (scale (1/4N^2)^2)
 *   I2integral ........................ src/integrals.c:211-251 (scale 1/4N^2)
 *   I3integral ........................ src/integrals.c:332-379 (scale (1/4N^2)^2)
 *   CdInt ............................. src/collision.c:74-168
 *   CcInt ............................. src/collision.c:209-254
 *   MagneticRotation .................. src/simulation.c:74-98
 *   CollisionStep (midpoint rule) ..... src/simulation.c:105-136
 *   r2c + Hermitian fill, c2c + Re .... src/simulation.c:224-283 (also 1229-1279)
 *   InhomStepPeriodic (non-MPI) ....... src/simulation.c:307-555 (:422-449, 452-453, 524-549)
 *   InhomStepDirichlet (non-MPI) ...... src/simulation.c:562-830 (collision skips the end cells, :715)
 *   InhomStepMaxwell (non-MPI) ........ src/simulation.c:837-1165 (Maxwell velocity grid, trace fluxes)
 *   SimulationHomogeneous ............. src/simulation.c:1219-1281
 * The reference's term order, pointwise-product operand order and scale
 * factors are kept, so this computes the reference's own (un-reassociated)
 * formulation: 68 + 96J + 24J^2 + 80JM FFTs per Cc+Cd evaluation.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   I1, I2 (custom psi), I2 (FourierI2), I3, I4 ..... test/data/I*_ref.dat
 *   CcInt / CdInt ................................... Cc*_ref.dat, Cd*_ref.dat
 *   CollisionStep ................................... WFnext*_ref.dat
 *   InhomStepPeriodic ............................... Wx_next_ref.dat, and frame 1
 *                                                     of data/Wevolv_inhom_periodic.dat
 *   SimulationHomogeneous ........................... frames 1..2 of data/Wevolv_hom.dat
 *   I1, I2, I3 structure ............................ brute-force direct convolution
 *                                                     at N = 4 (no FFT)
 *   MagneticRotation ................................ cross product closed form
 */
#include "qbf_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define CPLX(p, i) (&(p)[2 * (size_t)(i)])

void orc_zero_pad(int N, const double *f, double *ret)
{
	const int P = 2 * N;
	memset(ret, 0, sizeof(double) * 2 * (size_t)P * P);
	for (int l = 0; l < N / 2; l++) {
		memcpy(CPLX(ret, (size_t)P * l), CPLX(f, (size_t)N * l), sizeof(double) * N);
		memcpy(CPLX(ret, 3 * N / 2 + (size_t)P * l), CPLX(f, N / 2 + (size_t)N * l), sizeof(double) * N);
	}
	for (int l = N / 2; l < N; l++) {
		memcpy(CPLX(ret, (size_t)P * (N + l)), CPLX(f, (size_t)N * l), sizeof(double) * N);
		memcpy(CPLX(ret, 3 * N / 2 + (size_t)P * (N + l)), CPLX(f, N / 2 + (size_t)N * l), sizeof(double) * N);
	}
}

void orc_extract(int N, const double *f, double *ret)
{
	const int P = 2 * N;
	for (int l = 0; l < N / 2; l++) {
		memcpy(CPLX(ret, (size_t)N * l), CPLX(f, (size_t)P * l), sizeof(double) * N);
		memcpy(CPLX(ret, N / 2 + (size_t)N * l), CPLX(f, 3 * N / 2 + (size_t)P * l), sizeof(double) * N);
	}
	for (int l = N / 2; l < N; l++) {
		memcpy(CPLX(ret, (size_t)N * l), CPLX(f, (size_t)P * (N + l)), sizeof(double) * N);
		memcpy(CPLX(ret, N / 2 + (size_t)N * l), CPLX(f, 3 * N / 2 + (size_t)P * (N + l)), sizeof(double) * N);
	}
}

/* ret = f * ret (src/grid.c:152-164, operand order kept) */
static void mul_assign_cc(size_t n, const double *f, double *ret)
{
	for (size_t i = 0; i < n; i++) {
		const double re = ret[2 * i], im = ret[2 * i + 1];
		ret[2 * i] = f[2 * i] * re - f[2 * i + 1] * im;
		ret[2 * i + 1] = f[2 * i] * im + f[2 * i + 1] * re;
	}
}

/* ret = f * g, complex (src/grid.c:62-71, 92-101) */
static void mul_cc(size_t n, const double *f, const double *g, double *ret)
{
	for (size_t i = 0; i < n; i++) {
		ret[2 * i] = f[2 * i] * g[2 * i] - f[2 * i + 1] * g[2 * i + 1];
		ret[2 * i + 1] = f[2 * i] * g[2 * i + 1] + f[2 * i + 1] * g[2 * i];
	}
}

/* ret = f(real) * g(complex) (src/grid.c:47-55) */
static void mul_rc(size_t n, const double *f, const double *g, double *ret)
{
	for (size_t i = 0; i < n; i++) {
		ret[2 * i] = f[i] * g[2 * i];
		ret[2 * i + 1] = f[i] * g[2 * i + 1];
	}
}

/* ret *= f(real) (src/grid.c:137-145) */
static void mul_assign_rc(size_t n, const double *f, double *ret)
{
	for (size_t i = 0; i < n; i++) {
		ret[2 * i] *= f[i];
		ret[2 * i + 1] *= f[i];
	}
}

void orc_I1integral(int N, const double *f, const double *g, const double *h,
	unsigned num, const double *psiR1, const double *psiR2, double *I1)
{
	const int P = 2 * N;
	const size_t NN = (size_t)N * N, PP = (size_t)P * P;
	const double scale = (1.0 / (4.0 * N * N)) * (1.0 / (4.0 * N * N));
	double *f2F = malloc(sizeof(double) * 2 * PP), *h2F = malloc(sizeof(double) * 2 * PP);
	double *tmp = malloc(sizeof(double) * 2 * NN), *tmp2 = malloc(sizeof(double) * 2 * PP);
	orc_zero_pad(N, f, f2F); orc_fft2d(P, P, f2F, -1);
	orc_zero_pad(N, h, h2F); orc_fft2d(P, P, h2F, -1);
	memset(I1, 0, sizeof(double) * 2 * NN);
	for (unsigned j = 0; j < num; j++) {
		mul_rc(NN, &psiR2[j * NN], g, tmp);
		orc_zero_pad(N, tmp, tmp2);
		orc_fft2d(P, P, tmp2, -1);
		mul_assign_cc(PP, f2F, tmp2);
		orc_fft2d(P, P, tmp2, +1);
		mul_assign_rc(PP, &psiR1[j * PP], tmp2);
		orc_fft2d(P, P, tmp2, -1);
		mul_assign_cc(PP, h2F, tmp2);
		orc_fft2d(P, P, tmp2, +1);
		orc_extract(N, tmp2, tmp);
		for (size_t i = 0; i < NN; i++) {
			I1[2 * i] += scale * tmp[2 * i];
			I1[2 * i + 1] += scale * tmp[2 * i + 1];
		}
	}
	free(tmp2); free(tmp); free(h2F); free(f2F);
}

void orc_I2integral(int N, const double *f, const double *g, const double *h,
	unsigned num, const double *psiR1, const double *psiR2, double *I2)
{
	const int P = 2 * N;
	const size_t NN = (size_t)N * N, PP = (size_t)P * P;
	const double scale = 1.0 / (4.0 * N * N);
	double *h2F = malloc(sizeof(double) * 2 * PP);
	double *fp = malloc(sizeof(double) * 2 * NN), *gp = malloc(sizeof(double) * 2 * NN);
	double *F = malloc(sizeof(double) * 2 * PP), *G = malloc(sizeof(double) * 2 * PP);
	double *T = malloc(sizeof(double) * 2 * PP), *tmp = malloc(sizeof(double) * 2 * NN);
	orc_zero_pad(N, h, h2F); orc_fft2d(P, P, h2F, -1);
	memset(I2, 0, sizeof(double) * 2 * NN);
	for (unsigned j = 0; j < num; j++) {
		mul_rc(NN, &psiR1[j * NN], f, fp);
		mul_rc(NN, &psiR2[j * NN], g, gp);
		orc_zero_pad(N, fp, F);
		orc_zero_pad(N, gp, G);
		orc_fft2d(P, P, F, -1);
		orc_fft2d(P, P, G, -1);
		mul_cc(PP, F, G, T);
		mul_assign_cc(PP, h2F, T);
		orc_fft2d(P, P, T, +1);
		orc_extract(N, T, tmp);
		for (size_t i = 0; i < NN; i++) {
			I2[2 * i] += scale * tmp[2 * i];
			I2[2 * i + 1] += scale * tmp[2 * i + 1];
		}
	}
	free(tmp); free(T); free(G); free(F); free(gp); free(fp); free(h2F);
}

void orc_I3integral(int N, const double *f, const double *g, const double *h,
	unsigned num, const double *psiR1, const double *psiR2, double *I3)
{
	const int P = 2 * N;
	const size_t NN = (size_t)N * N, PP = (size_t)P * P;
	const double scale = (1.0 / (4.0 * N * N)) * (1.0 / (4.0 * N * N));
	double *g2F = malloc(sizeof(double) * 2 * PP);
	double *fp = malloc(sizeof(double) * 2 * NN), *hp = malloc(sizeof(double) * 2 * NN);
	double *F = malloc(sizeof(double) * 2 * PP), *H = malloc(sizeof(double) * 2 * PP);
	double *T = malloc(sizeof(double) * 2 * PP), *tmp = malloc(sizeof(double) * 2 * NN);
	orc_zero_pad(N, g, g2F); orc_fft2d(P, P, g2F, -1);
	memset(I3, 0, sizeof(double) * 2 * NN);
	for (unsigned j = 0; j < num; j++) {
		mul_cc(NN, &psiR1[2 * j * NN], f, fp);
		mul_cc(NN, &psiR1[2 * j * NN], h, hp);
		orc_zero_pad(N, fp, F);
		orc_zero_pad(N, hp, H);
		orc_fft2d(P, P, F, -1);
		orc_fft2d(P, P, H, -1);
		mul_cc(PP, g2F, H, T);
		orc_fft2d(P, P, T, +1);
		mul_assign_rc(PP, &psiR2[j * PP], T);
		orc_fft2d(P, P, T, -1);
		mul_assign_cc(PP, F, T);
		orc_fft2d(P, P, T, +1);
		orc_extract(N, T, tmp);
		for (size_t i = 0; i < NN; i++) {
			I3[2 * i] += scale * tmp[2 * i];
			I3[2 * i + 1] += scale * tmp[2 * i + 1];
		}
	}
	free(tmp); free(T); free(H); free(F); free(hp); free(fp); free(g2F);
}

void orc_CdInt(const orc_quad_t *q, const double *W, double *Cd)
{
	const int N = q->N;
	const size_t NN = (size_t)N * N;
	const double eta[4] = { 1, -1, -1, -1 };
	const double eta2[4] = { 2, -2, -2, -2 };
	double *Wt = malloc(sizeof(double) * 8 * NN), *tmp = malloc(sizeof(double) * 2 * NN);
	memset(Cd, 0, sizeof(double) * 8 * NN);
	for (size_t e = 0; e < 8 * NN; e++) Wt[e] = -W[e];
	Wt[0] += 1;
	for (int j = 0; j < 4; j++)
		for (int i = 0; i < 4; i++) {
			orc_I2integral(N, &W[2 * i * NN], &W[2 * i * NN], &Wt[2 * j * NN], q->numI2, q->I2_psiR1, q->I2_psiR2, tmp);
			for (size_t k = 0; k < 2 * NN; k++) Cd[2 * j * NN + k] += eta2[i] * tmp[k];
		}
	for (int j = 0; j < 4; j++)
		for (int i = 0; i < 4; i++) {
			orc_I3integral(N, &W[2 * i * NN], &W[2 * i * NN], &W[2 * j * NN], q->numI3, q->I3_psiR1, q->I3_psiR2, tmp);
			for (size_t k = 0; k < 2 * NN; k++) Cd[2 * j * NN + k] -= eta[i] * eta2[j] * tmp[k];
		}
	for (size_t k = 0; k < 2 * NN; k++) Wt[k] = 2 * W[k];
	Wt[0] -= 1;
	for (int j = 0; j < 4; j++) {
		orc_I1integral(N, Wt, &W[2 * j * NN], W, q->numI4, q->I4_psiR1, q->I4_psiR2, tmp);
		for (size_t k = 0; k < 2 * NN; k++) Cd[2 * j * NN + k] += eta2[j] * tmp[k];
	}
	for (int j = 0; j < 4; j++) {
		orc_I1integral(N, Wt, W, &W[2 * j * NN], q->numI4, q->I4_psiR1, q->I4_psiR2, tmp);
		for (size_t k = 0; k < 2 * NN; k++) Cd[2 * j * NN + k] += 2 * tmp[k];
	}
	for (int j = 0; j < 4; j++) {
		orc_I1integral(N, Wt, &W[2 * j * NN], &W[2 * j * NN], q->numI4, q->I4_psiR1, q->I4_psiR2, tmp);
		for (size_t k = 0; k < 2 * NN; k++) Cd[k] -= 2 * tmp[k];
	}
	free(tmp); free(Wt);
}

void orc_CcInt(const orc_quad_t *q, const double *W, double *Cc)
{
	const int N = q->N;
	const size_t NN = (size_t)N * N;
	double *f = malloc(sizeof(double) * 2 * NN), *I1 = malloc(sizeof(double) * 2 * NN);
	memset(Cc, 0, sizeof(double) * 8 * NN);
	for (size_t k = 0; k < 2 * NN; k++) f[k] = -4 * W[k];
	f[0] += 2;
	static const int pairs[3][2] = { { 2, 3 }, { 3, 1 }, { 1, 2 } };
	for (int c = 0; c < 3; c++) {
		const int a = pairs[c][0], b = pairs[c][1];
		double *out = &Cc[2 * (c + 1) * NN];
		orc_I1integral(N, f, &W[2 * a * NN], &W[2 * b * NN], q->numI1, q->I1_psiR1, q->I1_psiR2, out);
		orc_I1integral(N, f, &W[2 * b * NN], &W[2 * a * NN], q->numI1, q->I1_psiR1, q->I1_psiR2, I1);
		for (size_t k = 0; k < 2 * NN; k++) out[k] -= I1[k];
	}
	free(I1); free(f);
}

void orc_MagneticRotation(int N, const double *W, const double Bext[3], double *CB)
{
	const size_t NN = (size_t)N * N;
	const double *w1 = &W[2 * NN], *w2 = &W[4 * NN], *w3 = &W[6 * NN];
	memset(CB, 0, sizeof(double) * 8 * NN);
	for (size_t k = 0; k < 2 * NN; k++) {
		CB[2 * NN + k] = 2 * (Bext[1] * w3[k] - Bext[2] * w2[k]);
		CB[4 * NN + k] = 2 * (Bext[2] * w1[k] - Bext[0] * w3[k]);
		CB[6 * NN + k] = 2 * (Bext[0] * w2[k] - Bext[1] * w1[k]);
	}
}

void orc_CollisionStep(const orc_quad_t *q, const double *W, double dt, const double Bext[3], double *Wnext)
{
	const size_t n8 = 8 * (size_t)q->N * q->N;
	double *Cc = malloc(sizeof(double) * n8), *Cd = malloc(sizeof(double) * n8);
	double *CB = malloc(sizeof(double) * n8), *Wmid = malloc(sizeof(double) * n8);
	orc_CcInt(q, W, Cc);
	orc_CdInt(q, W, Cd);
	orc_MagneticRotation(q->N, W, Bext, CB);
	const double dth = 0.5 * dt;
	for (size_t k = 0; k < n8; k++) Wmid[k] = W[k] + dth * (Cc[k] + Cd[k] + CB[k]);
	orc_CcInt(q, Wmid, Cc);
	orc_CdInt(q, Wmid, Cd);
	orc_MagneticRotation(q->N, Wmid, Bext, CB);
	for (size_t k = 0; k < n8; k++) Wnext[k] = W[k] + dt * (Cc[k] + Cd[k] + CB[k]);
	free(Wmid); free(CB); free(Cd); free(Cc);
}

void orc_r2c_fill(int N, const double *in, double *out)
{
	const size_t NN = (size_t)N * N;
	double *X = malloc(sizeof(double) * 2 * NN);
	for (size_t i = 0; i < NN; i++) { X[2 * i] = in[i]; X[2 * i + 1] = 0; }
	orc_fft2d(N, N, X, -1);
	const double scale = 1.0 / ((double)N * N);
	for (int l = 0; l < N; l++) {
		const int lc = (N - l) & (N - 1);
		for (int k = 0; k <= N / 2; k++) {
			out[2 * (N * l + k)] = X[2 * (N * l + k)] * scale;
			out[2 * (N * l + k) + 1] = X[2 * (N * l + k) + 1] * scale;
		}
		for (int k = N / 2 + 1; k < N; k++) {
			const int kc = N - k;
			out[2 * (N * l + k)] = X[2 * (N * lc + kc)] * scale;
			out[2 * (N * l + k) + 1] = -(X[2 * (N * lc + kc) + 1] * scale);
		}
	}
	free(X);
}

void orc_c2r_real(int N, const double *in, double *out)
{
	const size_t NN = (size_t)N * N;
	double *X = malloc(sizeof(double) * 2 * NN);
	memcpy(X, in, sizeof(double) * 2 * NN);
	orc_fft2d(N, N, X, +1);
	for (size_t i = 0; i < NN; i++) out[i] = X[2 * i];
	free(X);
}

/* one cell: physical -> Fourier -> CollisionStep -> physical (src/simulation.c:224-283) */
static void fourier_collision_cell(const orc_quad_t *q, const double *Win, double dt, const double *Bext, double *Wout)
{
	const int N = q->N;
	const size_t NN = (size_t)N * N;
	double *WF = malloc(sizeof(double) * 8 * NN), *WF2 = malloc(sizeof(double) * 8 * NN);
	for (int c = 0; c < 4; c++) orc_r2c_fill(N, &Win[c * NN], &WF[2 * c * NN]);
	orc_CollisionStep(q, WF, dt, Bext, WF2);
	for (int c = 0; c < 4; c++) orc_c2r_real(N, &WF2[2 * c * NN], &Wout[c * NN]);
	free(WF2); free(WF);
}

static void half_transport(int N, unsigned numVol, double L, double h, double dt, const double *Win, double *Wout)
{
	const size_t NN = (size_t)N * N;
	const double scale = (2 * L) / N;
	double *Wx = malloc(sizeof(double) * numVol), *Wx1 = malloc(sizeof(double) * numVol);
	for (int j = 0; j < 4; j++)
		for (int l = 0; l < N; l++) {
			const double v = l < N / 2 ? scale * l : scale * (l - N);
			for (int k = 0; k < N; k++) {
				const size_t off = j * NN + k + (size_t)N * l;
				for (unsigned i = 0; i < numVol; i++) Wx[i] = Win[i * 4 * NN + off];
				orc_SlopeLimiterStepPeriodic(h, 0.5 * dt, v, numVol, Wx, Wx1);
				for (unsigned i = 0; i < numVol; i++) Wout[i * 4 * NN + off] = Wx1[i];
			}
		}
	free(Wx1); free(Wx);
}

void orc_InhomStepPeriodic(const orc_quad_t *q, unsigned numVol, double L, const double *W,
	double h, double dt, const double *Bext, double *Wnext)
{
	const int N = q->N;
	const size_t cell = 4 * (size_t)N * N;
	double *Wtmp = malloc(sizeof(double) * cell * numVol), *Wtmp2 = malloc(sizeof(double) * cell * numVol);
	half_transport(N, numVol, L, h, dt, W, Wtmp);
	#pragma omp parallel for schedule(dynamic, 1)
	for (int i = 0; i < (int)numVol; i++)
		fourier_collision_cell(q, &Wtmp[i * cell], dt, &Bext[3 * i], &Wtmp2[i * cell]);
	half_transport(N, numVol, L, h, dt, Wtmp2, Wnext);
	free(Wtmp2); free(Wtmp);
}

/* orc_InhomStepPeriodic evaluated on a subset: the collision runs only on the cells within two of
 * a listed cell (the limiter's stencil, src/finite_volume.c:74-100, reaches two cells to each side),
 * so Wnext is exact (bit for bit the full step's) on the listed cells and undefined elsewhere. */
void orc_InhomStepPeriodicCells(const orc_quad_t *q, unsigned numVol, double L, const double *W,
	double h, double dt, const double *Bext, unsigned nlist, const unsigned *list, double *Wnext)
{
	const int N = q->N;
	const size_t cell = 4 * (size_t)N * N;
	double *Wtmp = malloc(sizeof(double) * cell * numVol), *Wtmp2 = malloc(sizeof(double) * cell * numVol);
	char *need = calloc(numVol, 1);
	for (unsigned j = 0; j < nlist; j++)
		for (int d = -2; d <= 2; d++) need[((int)list[j] + d + 2 * (int)numVol) % (int)numVol] = 1;
	half_transport(N, numVol, L, h, dt, W, Wtmp);
	memcpy(Wtmp2, Wtmp, sizeof(double) * cell * numVol);
	#pragma omp parallel for schedule(dynamic, 1)
	for (int i = 0; i < (int)numVol; i++)
		if (need[i]) fourier_collision_cell(q, &Wtmp[i * cell], dt, &Bext[3 * i], &Wtmp2[i * cell]);
	half_transport(N, numVol, L, h, dt, Wtmp2, Wnext);
	free(need); free(Wtmp2); free(Wtmp);
}

static void half_transport_dirichlet(int N, unsigned numVol, double L, double h, double dt, const double *Win,
	double *Wout)
{
	const size_t NN = (size_t)N * N;
	const double scale = (2 * L) / N;
	double *Wx = malloc(sizeof(double) * numVol), *Wx1 = malloc(sizeof(double) * numVol);
	for (int j = 0; j < 4; j++)
		for (int l = 0; l < N; l++) {
			const double v = l < N / 2 ? scale * l : scale * (l - N);
			for (int k = 0; k < N; k++) {
				const size_t off = j * NN + k + (size_t)N * l;
				for (unsigned i = 0; i < numVol; i++) Wx[i] = Win[i * 4 * NN + off];
				orc_SlopeLimiterStepDirichlet(h, 0.5 * dt, v, numVol, Wx, Wx1);
				for (unsigned i = 0; i < numVol; i++) Wout[i * 4 * NN + off] = Wx1[i];
			}
		}
	free(Wx1); free(Wx);
}

/* half transport with Dirichlet ends, collision of the cells 1 .. numVol-2 (the end cells keep
 * their transported values), half transport */
void orc_InhomStepDirichlet(const orc_quad_t *q, unsigned numVol, double L, const double *W,
	double h, double dt, const double *Bext, double *Wnext)
{
	const int N = q->N;
	const size_t cell = 4 * (size_t)N * N;
	double *Wtmp = malloc(sizeof(double) * cell * numVol), *Wtmp2 = malloc(sizeof(double) * cell * numVol);
	half_transport_dirichlet(N, numVol, L, h, dt, W, Wtmp);
	memcpy(Wtmp2, Wtmp, sizeof(double) * cell * numVol);
	#pragma omp parallel for schedule(dynamic, 1)
	for (int i = 1; i < (int)numVol - 1; i++)
		fourier_collision_cell(q, &Wtmp[i * cell], dt, &Bext[3 * i], &Wtmp2[i * cell]);
	half_transport_dirichlet(N, numVol, L, h, dt, Wtmp2, Wnext);
	free(Wtmp2); free(Wtmp);
}

/* Maxwell velocity grid (src/simulation.c:155-161, :837-856): the zero x-velocity row is left
 * out and the -L row appears twice, as -L (row N/2) and +L (row N/2-1); x velocity on the slow index */
static void to_maxwell_grid(int N, const double *c, double *m)
{
	const size_t half = (size_t)N * N / 2;
	memcpy(m, &c[N], sizeof(double) * half);
	memcpy(&m[half], &c[half], sizeof(double) * half);
}

static void from_maxwell_grid(int N, const double *m, const double *v0, double *c)
{
	const size_t half = (size_t)N * N / 2;
	memcpy(c, v0, sizeof(double) * N);
	memcpy(&c[N], m, sizeof(double) * (half - N));
	memcpy(&c[half], &m[half], sizeof(double) * half);
}

/* half transport with Maxwell reflection at both ends (src/simulation.c:905-1051): outgoing trace
 * fluxes of the end cells, the limiter on the Maxwell grid per component, zero x-velocity row kept */
static void half_transport_maxwell(int N, unsigned numVol, const double *vxM, double h, double dt, double lambda,
	const double *WLeff, const double *WReff, const double *Win, double *Wout)
{
	const size_t NN = (size_t)N * N, cell = 4 * NN;
	double *tr = malloc(sizeof(double) * NN);
	to_maxwell_grid(N, &Win[0], tr);
	const double fluxL = orc_CalculateLeftwardFlux((unsigned)NN, vxM, tr);
	to_maxwell_grid(N, &Win[(numVol - 1) * cell], tr);
	const double fluxR = orc_CalculateRightwardFlux((unsigned)NN, vxM, tr);
	double *Wx = malloc(sizeof(double) * NN * numVol), *Wx1 = malloc(sizeof(double) * NN * numVol);
	for (int j = 0; j < 4; j++) {
		for (unsigned i = 0; i < numVol; i++) to_maxwell_grid(N, &Win[i * cell + j * NN], &Wx[i * NN]);
		orc_SlopeLimiterStepMaxwell(h, 0.5 * dt, (unsigned)NN, vxM, numVol, lambda, fluxL, fluxR,
			&WLeff[j * NN], &WReff[j * NN], Wx, Wx1);
		for (unsigned i = 0; i < numVol; i++)
			from_maxwell_grid(N, &Wx1[i * NN], &Win[i * cell + j * NN], &Wout[i * cell + j * NN]);
	}
	free(Wx1); free(Wx); free(tr);
}

/* Maxwell boundary states on the Maxwell grid, normalised by their incoming trace flux
 * (src/simulation.c:882-904), and the x velocities of that grid */
static void maxwell_setup(int N, double L, const double *WMaxwL, const double *WMaxwR, double *vxM, double *WL, double *WR)
{
	const size_t NN = (size_t)N * N, cell = 4 * NN;
	const double scale = (2 * L) / N;
	for (int l = 0; l < N; l++)
		for (int k = 0; k < N; k++) vxM[k + (size_t)N * l] = l < N / 2 ? scale * (l + 1) : scale * (l - N);
	for (int j = 0; j < 4; j++) {
		to_maxwell_grid(N, &WMaxwL[j * NN], &WL[j * NN]);
		to_maxwell_grid(N, &WMaxwR[j * NN], &WR[j * NN]);
	}
	const double normL = 1.0 / orc_CalculateRightwardFlux((unsigned)NN, vxM, WL);
	const double normR = 1.0 / orc_CalculateLeftwardFlux((unsigned)NN, vxM, WR);
	for (size_t i = 0; i < cell; i++) {
		WL[i] *= normL;
		WR[i] *= normR;
	}
}

/* one half transport (dt / 2) of the Maxwell step, as orc_InhomStepMaxwell applies it */
void orc_TransportMaxwell(int N, unsigned numVol, double L, const double *W, double h, double dt, double lambda,
	const double *WMaxwL, const double *WMaxwR, double *Wout)
{
	const size_t NN = (size_t)N * N, cell = 4 * NN;
	double *vxM = malloc(sizeof(double) * NN), *WL = malloc(sizeof(double) * cell), *WR = malloc(sizeof(double) * cell);
	maxwell_setup(N, L, WMaxwL, WMaxwR, vxM, WL, WR);
	half_transport_maxwell(N, numVol, vxM, h, dt, lambda, WL, WR, W, Wout);
	free(WR); free(WL); free(vxM);
}

/* Maxwell boundary step (src/simulation.c:882-1165): boundary states on the Maxwell grid,
 * normalised by their incoming trace flux; half transport, collision of every cell, half transport */
void orc_InhomStepMaxwell(const orc_quad_t *q, unsigned numVol, double L, const double *W, double h, double dt,
	double lambda, const double *WMaxwL, const double *WMaxwR, const double *Bext, double *Wnext)
{
	const int N = q->N;
	const size_t NN = (size_t)N * N, cell = 4 * NN;
	double *vxM = malloc(sizeof(double) * NN), *WL = malloc(sizeof(double) * cell), *WR = malloc(sizeof(double) * cell);
	maxwell_setup(N, L, WMaxwL, WMaxwR, vxM, WL, WR);
	double *Wtmp = malloc(sizeof(double) * cell * numVol), *Wtmp2 = malloc(sizeof(double) * cell * numVol);
	half_transport_maxwell(N, numVol, vxM, h, dt, lambda, WL, WR, W, Wtmp);
	#pragma omp parallel for schedule(dynamic, 1)
	for (int i = 0; i < (int)numVol; i++)
		fourier_collision_cell(q, &Wtmp[i * cell], dt, &Bext[3 * i], &Wtmp2[i * cell]);
	half_transport_maxwell(N, numVol, vxM, h, dt, lambda, WL, WR, Wtmp2, Wnext);
	free(Wtmp2); free(Wtmp); free(WR); free(WL); free(vxM);
}

void orc_SimulationHomogeneous(const orc_quad_t *q, const double *W0, double dt, unsigned numsteps,
	const double Bext[3], double *Wevolv)
{
	const size_t cell = 4 * (size_t)q->N * q->N;
	memcpy(Wevolv, W0, sizeof(double) * cell);
	for (unsigned it = 1; it < numsteps; it++)
		fourier_collision_cell(q, &Wevolv[(it - 1) * cell], dt, Bext, &Wevolv[it * cell]);
}

void orc_moments(int N, double L, const double *W, double *mom)
{
	const size_t NN = (size_t)N * N;
	const double scale = (2 * L) / N, dv = scale * scale;
	for (int i = 0; i < 7; i++) mom[i] = 0;
	for (int l = 0; l < N; l++) {
		const double vx = l < N / 2 ? scale * l : scale * (l - N);
		for (int k = 0; k < N; k++) {
			const double vy = k < N / 2 ? scale * k : scale * (k - N);
			const size_t e = k + (size_t)N * l;
			for (int c = 0; c < 4; c++) mom[c] += W[c * NN + e];
			mom[4] += vx * W[e];
			mom[5] += vy * W[e];
			mom[6] += 0.5 * (vx * vx + vy * vy) * W[e];
		}
	}
	for (int i = 0; i < 7; i++) mom[i] *= dv;
}
