/* Host paths of the drop-in layer and the oracle, run under AddressSanitizer and UBSan (gcc, no
 * device: the device API comes from engine_stub.c).  Each check prints one line; any sanitizer report
 * or failed check ends the program with a non-zero status.
 *
 *   ReadData / WriteData: the -1 (cannot open) and -2 (size mismatch) codes, a round trip, append
 *   the FFTW subset (r2c plans; execute without a device leaves the output alone)
 *   GridZeroPad / GridExtract round trip, MultiplyPointwise* against direct products
 *   FourierI1..I4 and LegendreGaussQuad (tables filled, weights sum to b - a), QuadI*_Delete
 *   the host limiters (Dirichlet, Maxwell, trace fluxes) against the oracle's, bit for bit
 *   the device-backed entry points without a device: CcInt returns with its output untouched
 *   the oracle at N = 8: FFT, integrals, Cc, Cd, collision, periodic / Dirichlet / Maxwell steps,
 *   the homogeneous simulation, moments
 * Usage: host_paths <scratch directory> */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "collision.h"
#include "finite_volume.h"
#include "grid.h"
#include "quadrature.h"
#include "qbf_oracle.h"
#include "util.h"
#include "wigner.h"

static int fails = 0;

static void check(int ok, const char *what)
{
	printf("%s: %s\n", ok ? "ok" : "FAILED", what);
	if (!ok) fails++;
}

static unsigned long long rng_state = 12345;
static double urand(void)
{
	rng_state = rng_state * 6364136223846793005ULL + 1442695040888963407ULL;
	return (double)(rng_state >> 11) / 9007199254740992.0;
}

static void io_paths(const char *dir)
{
	char good[1024], missing[1024];
	snprintf(good, sizeof good, "%s/io_test.dat", dir);
	snprintf(missing, sizeof missing, "%s/does_not_exist.dat", dir);
	double a[16], b[16];
	for (int i = 0; i < 16; i++) a[i] = urand();
	check(WriteData(good, a, sizeof(double), 16, false) == 0, "WriteData writes a file");
	check(ReadData(good, b, sizeof(double), 16) == 0 && memcmp(a, b, sizeof a) == 0, "ReadData reads it back");
	check(ReadData(good, b, sizeof(double), 15) == -2, "ReadData: size mismatch is -2");
	check(ReadData(missing, b, sizeof(double), 16) == -1, "ReadData: missing file is -1");
	check(WriteData(good, a, sizeof(double), 16, true) == 0, "WriteData appends");
	double c[32];
	check(ReadData(good, c, sizeof(double), 32) == 0 && memcmp(c + 16, a, sizeof a) == 0, "appended data read back");
	char nodir[1024];
	snprintf(nodir, sizeof nodir, "%s/no/such/dir.dat", dir);
	check(WriteData(nodir, a, sizeof(double), 16, false) == -1, "WriteData: unwritable path is -1");
}

static void fftw_paths(void)
{
	/* the subset's transform runs on the device: without one, fftw_execute reports and leaves the
	 * output alone; plans of another size are refused */
	const int N = N_GRID;
	double *in = fftw_malloc(sizeof(double) * N * N);
	fftw_complex *out = fftw_malloc(sizeof(fftw_complex) * N * (N / 2 + 1));
	for (int i = 0; i < N * N; i++) in[i] = urand() - 0.5;
	memset(out, 0x5a, sizeof(fftw_complex) * N * (N / 2 + 1));
	fftw_plan p = fftw_plan_dft_r2c_2d(N, N, in, out, FFTW_ESTIMATE);
	check(p != NULL, "fftw_plan_dft_r2c_2d at N_GRID");
	fftw_execute(p);
	fftw_destroy_plan(p);
	const unsigned char *b = (const unsigned char *)out;
	int untouched = 1;
	for (size_t i = 0; i < sizeof(fftw_complex) * N * (N / 2 + 1); i++) untouched &= b[i] == 0x5a;
	check(untouched, "fftw_execute without a device leaves the output alone");
	check(fftw_plan_dft_r2c_2d(N / 2, N, in, out, FFTW_ESTIMATE) == NULL, "fftw_plan_dft_r2c_2d refuses other sizes");
	fftw_free(in);
	fftw_free(out);
}

static void grid_paths(void)
{
	cgrid_t *f = malloc(sizeof(cgrid_t)), *g = malloc(sizeof(cgrid_t)), *r = malloc(sizeof(cgrid_t));
	cgrid2_t *pd = malloc(sizeof(cgrid2_t));
	grid_t *x = malloc(sizeof(grid_t));
	for (int i = 0; i < N_GRID * N_GRID; i++) {
		f->data[i][0] = urand();
		f->data[i][1] = urand();
		g->data[i][0] = urand();
		g->data[i][1] = urand();
		x->data[i] = urand();
	}
	GridZeroPad(f, pd);
	GridExtract(pd, r);
	check(memcmp(f, r, sizeof(cgrid_t)) == 0, "GridExtract(GridZeroPad(f)) = f");
	int zeros = 0;
	for (int i = 0; i < 4 * N_GRID * N_GRID; i++) zeros += pd->data[i][0] == 0 && pd->data[i][1] == 0;
	check(zeros == 3 * N_GRID * N_GRID, "GridZeroPad leaves three quarters zero");
	MultiplyPointwiseCC(f, g, r);
	double e = 0;
	for (int i = 0; i < N_GRID * N_GRID; i++) {
		e = fmax(e, fabs(r->data[i][0] - (f->data[i][0] * g->data[i][0] - f->data[i][1] * g->data[i][1])));
		e = fmax(e, fabs(r->data[i][1] - (f->data[i][0] * g->data[i][1] + f->data[i][1] * g->data[i][0])));
	}
	check(e < 1e-15, "MultiplyPointwiseCC");
	MultiplyPointwiseRC(x, g, r);
	e = 0;
	for (int i = 0; i < N_GRID * N_GRID; i++) e = fmax(e, fabs(r->data[i][0] - x->data[i] * g->data[i][0]));
	check(e == 0, "MultiplyPointwiseRC");
	free(f);
	free(g);
	free(r);
	free(pd);
	free(x);
}

static void quadrature_paths(void)
{
	double x[16], w[16];
	LegendreGaussQuad(16, -2.0, 3.0, x, w);
	double s = 0;
	for (int i = 0; i < 16; i++) s += w[i];
	check(fabs(s - 5.0) < 1e-13 && x[0] > x[15], "LegendreGaussQuad: weights sum to b - a, nodes descending");
	quadI1_t q1, q4;
	quadI2_t q2;
	quadI3_t q3;
	FourierI1(4, 6.0, 3.75, &q1);
	FourierI2(4, 6.0, 3.75, &q2);
	FourierI3(4, 6.0, 3.75, 2, &q3);
	FourierI4(4, 6.0, 3.75, &q4);
	check(q1.num == 16 && q2.num == 4 && q3.num == 8 && q4.num == 4, "FourierI1..I4 term counts");
	int finite = 1;
	for (unsigned i = 0; i < q3.num; i++)
		for (int k = 0; k < 4 * N_GRID * N_GRID; k++) finite &= isfinite(q3.psiR2[i].data[k]);
	check(finite, "FourierI3 tables finite");
	/* the device-backed collision without a device: an error is reported and the output left alone */
	wignerF_t *W = calloc(1, sizeof(wignerF_t)), *Cc = malloc(sizeof(wignerF_t));
	memset(Cc, 0x5a, sizeof(wignerF_t));
	unsigned char sentinel[64];
	memset(sentinel, 0x5a, sizeof sentinel);
	CcInterm_t ci;
	CcInterm_Create(&ci);
	CcInt(W, &q1, &ci, Cc);
	CcInterm_Delete(&ci);
	check(memcmp(Cc, sentinel, sizeof sentinel) == 0 && strstr(qbf_last_error(), "device") != NULL,
		"CcInt without a device: error reported, output untouched");
	free(W);
	free(Cc);
	QuadI1_Delete(&q1);
	QuadI2_Delete(&q2);
	QuadI3_Delete(&q3);
	QuadI1_Delete(&q4);
}

static void limiter_paths(void)
{
	enum { NX = 37, M = 16 };
	double U[NX], a[NX], b[NX], A[M], Um[M * NX], UL[M], UR[M], c[M * NX], d[M * NX];
	for (int i = 0; i < NX; i++) U[i] = urand();
	SlopeLimiterStepDirichlet(0.08, 0.02, -1.2, NX, U, a);
	orc_SlopeLimiterStepDirichlet(0.08, 0.02, -1.2, NX, U, b);
	check(memcmp(a, b, sizeof a) == 0, "SlopeLimiterStepDirichlet = oracle, bit for bit");
	for (int i = 0; i < M; i++) {
		A[i] = (i < M / 2 ? i + 1 : i - M) * 0.75;
		UL[i] = urand();
		UR[i] = urand();
	}
	for (int i = 0; i < M * NX; i++) Um[i] = urand();
	check(CalculateRightwardFlux(M, A, UL) == orc_CalculateRightwardFlux(M, A, UL) &&
		CalculateLeftwardFlux(M, A, UR) == orc_CalculateLeftwardFlux(M, A, UR), "trace fluxes = oracle");
	SlopeLimiterStepMaxwell(0.05, 0.01, M, A, NX, 0.3, 0.7, 0.9, UL, UR, Um, c);
	orc_SlopeLimiterStepMaxwell(0.05, 0.01, M, A, NX, 0.3, 0.7, 0.9, UL, UR, Um, d);
	check(memcmp(c, d, sizeof c) == 0, "SlopeLimiterStepMaxwell = oracle, bit for bit");
}

static int all_finite(const double *x, size_t n)
{
	for (size_t i = 0; i < n; i++)
		if (!isfinite(x[i])) return 0;
	return 1;
}

static void oracle_paths(void)
{
	const int N = 8, P = 16, J = 4, M = 2, NV = 4;
	const size_t nn = (size_t)N * N, pp = (size_t)P * P;
	double *t1a = malloc(sizeof(double) * J * J * pp), *t1b = malloc(sizeof(double) * J * J * nn);
	double *t2a = malloc(sizeof(double) * J * nn), *t2b = malloc(sizeof(double) * J * nn);
	double *t3a = malloc(sizeof(double) * 2 * J * M * nn), *t3b = malloc(sizeof(double) * J * M * pp);
	double *t4a = malloc(sizeof(double) * J * pp), *t4b = malloc(sizeof(double) * J * nn);
	orc_FourierI1(N, J, 6.0, 3.75, t1a, t1b);
	orc_FourierI2(N, J, 6.0, 3.75, t2a, t2b);
	orc_FourierI3(N, J, 6.0, 3.75, M, t3a, t3b);
	orc_FourierI4(N, J, 6.0, 3.75, t4a, t4b);
	orc_quad_t q = {N, J * J, t1a, t1b, J, t2a, t2b, J * M, t3a, t3b, J, t4a, t4b};
	double *Wp = malloc(sizeof(double) * NV * 4 * nn), *WF = malloc(sizeof(double) * 2 * 4 * nn);
	for (int c = 0; c < NV; c++)
		for (size_t i = 0; i < nn; i++) {
			const double w0 = 0.3 + 0.2 * urand();
			Wp[(c * 4 + 0) * nn + i] = w0;
			for (int k = 1; k < 4; k++) Wp[(c * 4 + k) * nn + i] = 0.2 * w0 * (urand() - 0.5);
		}
	for (int k = 0; k < 4; k++) orc_r2c_fill(N, Wp + k * nn, WF + 2 * k * nn);
	double *out = malloc(sizeof(double) * 2 * 4 * nn), *out2 = malloc(sizeof(double) * 2 * 4 * nn);
	orc_I1integral(N, WF, WF + 2 * nn, WF + 4 * nn, J * J, t1a, t1b, out);
	orc_I2integral(N, WF, WF + 2 * nn, WF + 4 * nn, J, t2a, t2b, out + 2 * nn);
	orc_I3integral(N, WF, WF + 2 * nn, WF + 4 * nn, J * M, t3a, t3b, out + 4 * nn);
	check(all_finite(out, 6 * nn), "oracle I1, I2, I3 at N = 8");
	orc_CcInt(&q, WF, out);
	orc_CdInt(&q, WF, out2);
	check(all_finite(out, 8 * nn) && all_finite(out2, 8 * nn), "oracle Cc, Cd at N = 8");
	const double B[3] = {0.1, -0.2, 0.3};
	orc_CollisionStep(&q, WF, 1e-3, B, out);
	check(all_finite(out, 8 * nn), "oracle collision step");
	double *Wn = malloc(sizeof(double) * NV * 4 * nn), *Bx = calloc(3 * NV, sizeof(double));
	orc_InhomStepPeriodic(&q, NV, 6.0, Wp, 0.25, 1e-3, Bx, Wn);
	check(all_finite(Wn, NV * 4 * nn), "oracle periodic Strang step");
	orc_InhomStepDirichlet(&q, NV, 6.0, Wp, 0.25, 1e-3, Bx, Wn);
	check(all_finite(Wn, NV * 4 * nn), "oracle Dirichlet Strang step");
	orc_InhomStepMaxwell(&q, NV, 6.0, Wp, 0.25, 1e-3, 0.4, Wp, Wp + 4 * nn, Bx, Wn);
	check(all_finite(Wn, NV * 4 * nn), "oracle Maxwell Strang step");
	double *ev = malloc(sizeof(double) * 3 * 2 * 4 * nn);
	orc_SimulationHomogeneous(&q, WF, 1e-3, 3, B, ev);
	check(all_finite(ev, 3 * 8 * nn), "oracle homogeneous simulation");
	double mom[7];
	orc_moments(N, 6.0, Wp, mom);
	check(all_finite(mom, 7) && mom[0] > 0, "oracle moments");
	free(t1a); free(t1b); free(t2a); free(t2b); free(t3a); free(t3b); free(t4a); free(t4b);
	free(Wp); free(WF); free(out); free(out2); free(Wn); free(Bx); free(ev);
}

int main(int argc, char **argv)
{
	if (argc < 2) {
		fprintf(stderr, "usage: %s <scratch directory>\n", argv[0]);
		return 2;
	}
	io_paths(argv[1]);
	fftw_paths();
	grid_paths();
	quadrature_paths();
	limiter_paths();
	oracle_paths();
	fftw_cleanup();
	printf("%d failed\n", fails);
	return fails ? 1 : 0;
}
