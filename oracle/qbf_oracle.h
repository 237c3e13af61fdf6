/*
 * qbf_oracle.h -- CPU oracle for the collision-plus-transport step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (cmendl/inhom-quantum-boltzmann-fourier-c), written from the
 * reference's maths with the reference's term order and scale factors, but
 * with the velocity grid size N as a runtime argument and with its own FFT
 * (radix-2, written from the DFT definition) in place of FFTW.  It is the
 * checker for the HIP path: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  Nothing in the product links it.
 *
 * Conventions (same as the reference):
 *   - real N x N grid index k + N*l, l = slow (v_x), k = fast (v_y)
 *     (src/simulation.c:156-160);
 *   - complex values are interleaved double pairs (fftw_complex);
 *   - padded grid P = 2N, frequency axis [0..N-1, -N..-1] (src/quadrature.c:167-170);
 *   - forward FFT = exp(-2 pi i jk/n), backward = exp(+2 pi i jk/n), both
 *     unnormalised (FFTW_FORWARD / FFTW_BACKWARD);
 *   - a Wigner state is 4 Pauli components stored component-major
 *     (include/wigner.h:35-50).
 *
 * Parity status per function is recorded in the file header of qbf_oracle.c.
 */
#ifndef QBF_ORACLE_H
#define QBF_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- FFT (own implementation, power-of-two sizes) ---- */
void orc_fft1d(int n, double *x, int sign);                 /* in place, complex interleaved */
void orc_fft2d(int n0, int n1, double *x, int sign);        /* [n0][n1] row-major, in place */

/* ---- grid helpers (src/grid.c:171-210) ---- */
void orc_zero_pad(int N, const double *f, double *ret);     /* N x N complex -> 2N x 2N complex */
void orc_extract(int N, const double *f, double *ret);      /* 2N x 2N complex -> N x N complex */

/* ---- quadrature (src/quadrature.c) ----
 * Tables are caller-allocated with the sizes below (P = 2N):
 *   I1: psiR1[J*J][P*P] real, psiR2[J*J][N*N] real
 *   I2: psiR1[J][N*N] real,   psiR2[J][N*N] real
 *   I3: psiR1[J*M][N*N] complex, psiR2[J*M][P*P] real
 *   I4: psiR1[J][P*P] real,   psiR2[J][N*N] real            */
void orc_LegendreGaussQuad(unsigned M, double a, double b, double *x, double *w);
void orc_FourierI1(int N, unsigned J, double L, double R, double *psiR1, double *psiR2);
void orc_FourierI2(int N, unsigned J, double L, double R, double *psiR1, double *psiR2);
void orc_FourierI3(int N, unsigned J, double L, double R, unsigned M, double *psiR1, double *psiR2);
void orc_FourierI4(int N, unsigned J, double L, double R, double *psiR1, double *psiR2);

/* quadrature bundle used by the collision operator */
typedef struct {
	int N;
	unsigned numI1; const double *I1_psiR1, *I1_psiR2;
	unsigned numI2; const double *I2_psiR1, *I2_psiR2;
	unsigned numI3; const double *I3_psiR1, *I3_psiR2;
	unsigned numI4; const double *I4_psiR1, *I4_psiR2;
} orc_quad_t;

/* ---- integrals (src/integrals.c) ---- */
void orc_I1integral(int N, const double *f, const double *g, const double *h,
	unsigned num, const double *psiR1, const double *psiR2, double *I1);
void orc_I2integral(int N, const double *f, const double *g, const double *h,
	unsigned num, const double *psiR1, const double *psiR2, double *I2);
void orc_I3integral(int N, const double *f, const double *g, const double *h,
	unsigned num, const double *psiR1, const double *psiR2, double *I3);

/* ---- collision operators (src/collision.c), W/Cc/Cd: 4 x N*N complex ---- */
void orc_CcInt(const orc_quad_t *q, const double *W, double *Cc);
void orc_CdInt(const orc_quad_t *q, const double *W, double *Cd);

/* ---- time stepping (src/simulation.c) ---- */
void orc_MagneticRotation(int N, const double *W, const double Bext[3], double *CB);
void orc_CollisionStep(const orc_quad_t *q, const double *W, double dt, const double Bext[3], double *Wnext);
/* physical real N x N -> Fourier N x N complex (r2c, 1/N^2, Hermitian fill) */
void orc_r2c_fill(int N, const double *in, double *out);
/* Fourier N x N complex -> physical: unnormalised backward FFT, real part */
void orc_c2r_real(int N, const double *in, double *out);
void orc_InhomStepPeriodic(const orc_quad_t *q, unsigned numVol, double L, const double *W,
	double h, double dt, const double *Bext, double *Wnext);
void orc_InhomStepPeriodicCells(const orc_quad_t *q, unsigned numVol, double L, const double *W,
	double h, double dt, const double *Bext, unsigned nlist, const unsigned *list, double *Wnext);
void orc_InhomStepDirichlet(const orc_quad_t *q, unsigned numVol, double L, const double *W,
	double h, double dt, const double *Bext, double *Wnext);
void orc_TransportMaxwell(int N, unsigned numVol, double L, const double *W, double h, double dt, double lambda,
	const double *WMaxwL, const double *WMaxwR, double *Wout);
void orc_InhomStepMaxwell(const orc_quad_t *q, unsigned numVol, double L, const double *W, double h, double dt,
	double lambda, const double *WMaxwL, const double *WMaxwR, const double *Bext, double *Wnext);
void orc_SimulationHomogeneous(const orc_quad_t *q, const double *W0, double dt, unsigned numsteps,
	const double Bext[3], double *Wevolv);

/* ---- finite volume transport (src/finite_volume.c) ---- */
void orc_SlopeLimiterStepPeriodic(double h, double dt, double A, unsigned N, const double *Un, double *Un1);
void orc_SlopeLimiterStepDirichlet(double h, double dt, double A, unsigned N, const double *Un, double *Un1);
double orc_CalculateRightwardFlux(unsigned m, const double *A, const double *U);
double orc_CalculateLeftwardFlux(unsigned m, const double *A, const double *U);
void orc_SlopeLimiterStepMaxwell(double h, double dt, unsigned m, const double *A, unsigned N,
	double lambda, double fluxL, double fluxR, const double *UmaxwL, const double *UmaxwR,
	const double *Un, double *Un1);

/* ---- diagnostics (visualization/wignerMoments.m:15-37) ----
 * mom[0..3] = dv * sum w_c, mom[4..5] = dv * sum v_x w0, v_y w0,
 * mom[6] = dv * sum |v|^2/2 w0, dv = (2L/N)^2, for one physical state. */
void orc_moments(int N, double L, const double *W, double *mom);

#ifdef __cplusplus
}
#endif
#endif
