/*
 * fftw3.h -- the subset of the FFTW3 API used by the reference's drivers and tests,
 * implemented by libqbf (no FFTW is installed on the target machines).
 *
 * Callers: src/main_hom.c:61,79,97,103,105 and src/main_inhom.c:315,450,487,490
 * (fftw_malloc / fftw_free / fftw_cleanup); test/*_test.c (fftw_plan_dft_r2c_2d,
 * fftw_execute, fftw_destroy_plan, e.g. test/Cc_test.c:104-106).
 * r2c plans of N_GRID x N_GRID execute on the GPU through the engine; other sizes
 * make fftw_plan_dft_r2c_2d return NULL.
 */
#ifndef QBF_FFTW3_SUBSET_H
#define QBF_FFTW3_SUBSET_H

#include <stddef.h>
#include <stdio.h>      /* the FFTW header provides stdio.h; callers rely on it */

#ifdef __cplusplus
extern "C" {
#endif

typedef double fftw_complex[2];
typedef struct qbf_fftw_plan_s *fftw_plan;

#define FFTW_FORWARD (-1)
#define FFTW_BACKWARD (+1)
#define FFTW_MEASURE (0U)
#define FFTW_ESTIMATE (1U << 6)

void *fftw_malloc(size_t n);
void fftw_free(void *p);
void fftw_cleanup(void);

/* out[n0][n1/2+1] = unnormalised forward DFT of the real in[n0][n1] */
fftw_plan fftw_plan_dft_r2c_2d(int n0, int n1, double *in, fftw_complex *out, unsigned flags);
void fftw_execute(const fftw_plan p);
void fftw_destroy_plan(fftw_plan p);

#ifdef __cplusplus
}
#endif
#endif
