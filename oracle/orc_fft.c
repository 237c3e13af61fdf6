/*
 * orc_fft.c -- radix-2 complex FFT for the oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Written from the DFT definition X[k] = sum_j x[j] exp(sign * 2 pi i j k / n),
 * unnormalised in both directions: sign = -1 is FFTW_FORWARD, sign = +1 is
 * FFTW_BACKWARD, which is the convention of every fftw_plan_dft_2d call in the
 * reference (src/integrals.c:48-51, src/simulation.c:175-176).  Iterative
 * decimation-in-time after a bit-reversal permutation; twiddles come from a
 * per-size table of cos/sin(2 pi k / n) evaluated directly (no recurrences).
 *
 * Parity: pinned by tests/test_oracle_fft.py against numpy.fft and a direct
 * O(n^2) DFT.
 */
#include "qbf_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MAX_LOG2 13

typedef struct {
	int n;
	unsigned *rev;
	double *cs;          /* cs[2k] = cos(2 pi k/n), cs[2k+1] = sin(2 pi k/n), k < n/2 */
} orc_plan_t;

static orc_plan_t g_plans[ORC_MAX_LOG2 + 1];

static int ilog2(int n)
{
	int l = 0;
	while ((1 << l) < n) l++;
	return l;
}

__attribute__((constructor))
static void orc_fft_init(void)
{
	for (int lg = 0; lg <= ORC_MAX_LOG2; lg++) {
		int n = 1 << lg;
		orc_plan_t *p = &g_plans[lg];
		p->n = n;
		p->rev = malloc(sizeof(unsigned) * n);
		p->cs = malloc(sizeof(double) * (n > 1 ? n : 2));
		for (int i = 0; i < n; i++) {
			unsigned r = 0;
			for (int b = 0; b < lg; b++)
				if (i & (1 << b)) r |= 1u << (lg - 1 - b);
			p->rev[i] = r;
		}
		for (int k = 0; k < n / 2; k++) {
			const double a = 2.0 * M_PI * (double)k / (double)n;
			p->cs[2 * k] = cos(a);
			p->cs[2 * k + 1] = sin(a);
		}
	}
}

void orc_fft1d(int n, double *x, int sign)
{
	const int lg = ilog2(n);
	if ((1 << lg) != n || lg > ORC_MAX_LOG2) abort();
	const orc_plan_t *p = &g_plans[lg];

	for (int i = 0; i < n; i++) {
		const unsigned r = p->rev[i];
		if ((unsigned)i < r) {
			double t0 = x[2 * i], t1 = x[2 * i + 1];
			x[2 * i] = x[2 * r]; x[2 * i + 1] = x[2 * r + 1];
			x[2 * r] = t0; x[2 * r + 1] = t1;
		}
	}
	const double sg = (sign < 0) ? -1.0 : 1.0;
	for (int m = 2; m <= n; m <<= 1) {
		const int half = m >> 1;
		const int step = n / m;
		for (int b = 0; b < n; b += m) {
			for (int j = 0; j < half; j++) {
				const double wr = p->cs[2 * j * step];
				const double wi = sg * p->cs[2 * j * step + 1];
				double *a = &x[2 * (b + j)];
				double *c = &x[2 * (b + j + half)];
				const double tr = wr * c[0] - wi * c[1];
				const double ti = wr * c[1] + wi * c[0];
				c[0] = a[0] - tr; c[1] = a[1] - ti;
				a[0] += tr;       a[1] += ti;
			}
		}
	}
}

void orc_fft2d(int n0, int n1, double *x, int sign)
{
	/* rows (contiguous, length n1) */
	for (int r = 0; r < n0; r++) orc_fft1d(n1, &x[2 * (size_t)r * n1], sign);
	/* columns (stride n1, length n0) */
	double *col = malloc(sizeof(double) * 2 * n0);
	for (int c = 0; c < n1; c++) {
		for (int r = 0; r < n0; r++) {
			col[2 * r] = x[2 * ((size_t)r * n1 + c)];
			col[2 * r + 1] = x[2 * ((size_t)r * n1 + c) + 1];
		}
		orc_fft1d(n0, col, sign);
		for (int r = 0; r < n0; r++) {
			x[2 * ((size_t)r * n1 + c)] = col[2 * r];
			x[2 * ((size_t)r * n1 + c) + 1] = col[2 * r + 1];
		}
	}
	free(col);
}
