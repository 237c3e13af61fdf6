/* The device API that the drop-in layer (csrc/qbf_compat.c) calls, answered without a device: every
 * call fails with "no device in this build".  It lets the host-only paths of the drop-in layer (binary
 * I/O, the FFTW subset, grid helpers, quadrature tables, host limiters) and the error paths of the
 * device-backed entry points be built and run under AddressSanitizer / UBSan with gcc alone. */
#include <stddef.h>

#include "qbf.h"

static const char *msg = "no device in this build";

const char *qbf_last_error(void) { return msg; }
int qbf_engine_create(int N, int precision, int maxCells, qbf_engine **out)
{
	(void)N; (void)precision; (void)maxCells;
	if (out) *out = NULL;
	return -2;
}
void qbf_engine_destroy(qbf_engine *e) { (void)e; }
void *qbf_malloc(size_t bytes) { (void)bytes; return NULL; }
void qbf_free(void *p) { (void)p; }
int qbf_h2d(qbf_engine *e, void *d, const void *s, size_t n) { (void)e; (void)d; (void)s; (void)n; return -2; }
int qbf_d2h(qbf_engine *e, void *d, const void *s, size_t n) { (void)e; (void)d; (void)s; (void)n; return -2; }
int qbf_set_i1_table(qbf_engine *e, int slot, unsigned num, const double *a, const double *b)
{ (void)e; (void)slot; (void)num; (void)a; (void)b; return -2; }
int qbf_set_i2_table(qbf_engine *e, unsigned num, const double *a, const double *b)
{ (void)e; (void)num; (void)a; (void)b; return -2; }
int qbf_set_i3_table(qbf_engine *e, unsigned num, const double *a, const double *b)
{ (void)e; (void)num; (void)a; (void)b; return -2; }
int qbf_collision(qbf_engine *e, const double *W, int n, int mode, double *o, const double *b, const double *wc,
	const double *B, double cf)
{ (void)e; (void)W; (void)n; (void)mode; (void)o; (void)b; (void)wc; (void)B; (void)cf; return -2; }
int qbf_collision_step(qbf_engine *e, const double *W, int n, double dt, const double *B, double *o)
{ (void)e; (void)W; (void)n; (void)dt; (void)B; (void)o; return -2; }
int qbf_phys_to_fourier(qbf_engine *e, const double *a, int n, double *b) { (void)e; (void)a; (void)n; (void)b; return -2; }
int qbf_fourier_to_phys(qbf_engine *e, const double *a, int n, double *b) { (void)e; (void)a; (void)n; (void)b; return -2; }
int qbf_transport_line_periodic(qbf_engine *e, const double *U, int n, double A, double h, double dt, double *o)
{ (void)e; (void)U; (void)n; (void)A; (void)h; (void)dt; (void)o; return -2; }
int qbf_inhom_step_periodic(qbf_engine *e, const double *W, int n, double L, double h, double dt, const double *B,
	double *o)
{ (void)e; (void)W; (void)n; (void)L; (void)h; (void)dt; (void)B; (void)o; return -2; }
int qbf_inhom_step_dirichlet(qbf_engine *e, const double *W, int n, double L, double h, double dt, const double *B,
	double *o)
{ (void)e; (void)W; (void)n; (void)L; (void)h; (void)dt; (void)B; (void)o; return -2; }
int qbf_inhom_step_maxwell(qbf_engine *e, const double *W, int n, double L, double h, double dt, double lam,
	const double *a, const double *b, const double *B, double *o)
{ (void)e; (void)W; (void)n; (void)L; (void)h; (void)dt; (void)lam; (void)a; (void)b; (void)B; (void)o; return -2; }
int qbf_simulate_periodic(qbf_engine *e, const double *W, int n, double L, double h, double dt, unsigned ns,
	unsigned st, const double *B, double *o)
{ (void)e; (void)W; (void)n; (void)L; (void)h; (void)dt; (void)ns; (void)st; (void)B; (void)o; return -2; }
int qbf_simulate_dirichlet(qbf_engine *e, const double *W, int n, double L, double h, double dt, unsigned ns,
	unsigned st, const double *B, double *o)
{ (void)e; (void)W; (void)n; (void)L; (void)h; (void)dt; (void)ns; (void)st; (void)B; (void)o; return -2; }
int qbf_simulate_maxwell(qbf_engine *e, const double *W, int n, double L, double h, double dt, double lam,
	const double *a, const double *b, unsigned ns, unsigned st, const double *B, double *o)
{
	(void)e; (void)W; (void)n; (void)L; (void)h; (void)dt; (void)lam; (void)a; (void)b; (void)ns; (void)st; (void)B;
	(void)o;
	return -2;
}
int qbf_I1(qbf_engine *e, int slot, const double *f, const double *g, const double *h, double *o)
{ (void)e; (void)slot; (void)f; (void)g; (void)h; (void)o; return -2; }
int qbf_I2(qbf_engine *e, const double *f, const double *g, const double *h, double *o)
{ (void)e; (void)f; (void)g; (void)h; (void)o; return -2; }
int qbf_I3(qbf_engine *e, const double *f, const double *g, const double *h, double *o)
{ (void)e; (void)f; (void)g; (void)h; (void)o; return -2; }
