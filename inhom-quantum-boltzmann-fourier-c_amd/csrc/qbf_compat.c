/*
 * qbf_compat.c -- the reference's C entry points (collision.h, integrals.h, quadrature.h,
 * simulation.h, finite_volume.h, grid.h, util.h and the fftw3.h subset its callers use),
 * implemented over the MI355X engine of include/qbf.h.
 *
 * Callers (src/main_hom.c, src/main_inhom.c, test/*.c of the reference) recompile
 * against the headers in include/ and link libqbf.so instead of the reference objects
 * and FFTW.  Semantics follow the reference: all arrays are host memory, every call is
 * synchronous, outputs never alias inputs (restrict_), errors are not returned (void
 * functions) but recorded for qbf_last_error() and reported on stderr.
 *
 * The collision work of every entry point runs on the GPU (one FP64 engine for
 * N = N_GRID, created on first use).  Quadrature tables are uploaded once per table
 * object: tables produced by FourierI1..I4 are cached by address until QuadI*_Delete;
 * tables filled by the caller after QuadI*_Allocate are re-validated by a content hash
 * on every call.
 */
#include "collision.h"
#include "finite_volume.h"
#include "qbf.h"
#include "quadrature.h"
#include "simulation.h"
#include "util.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NN_ (N_GRID * N_GRID)
#define PP_ (4 * N_GRID * N_GRID)


static void report(const char *where)
{
	fprintf(stderr, "libqbf: %s failed: %s\n", where, qbf_last_error());
}

/* ------------------------------------------------------------------ engine and device buffers */
static qbf_engine *g_eng;
static int g_cells;                /* maxCells of g_eng */

enum { SLOT_I1, SLOT_I4, SLOT_I2, SLOT_I3, NSLOT };
typedef struct { const void *p1, *p2; unsigned num; uint64_t hash; int valid; } slot_key_t;
static slot_key_t g_slot[NSLOT];

static void invalidate_slots(void) { memset(g_slot, 0, sizeof(g_slot)); }

static qbf_engine *engine(int ncell)
{
	if (g_eng && ncell <= g_cells) return g_eng;
	if (g_eng) qbf_engine_destroy(g_eng);
	g_eng = NULL;
	invalidate_slots();
	int cells = ncell > 16 ? ncell : 16;
	if (qbf_engine_create(N_GRID, QBF_FP64, cells, &g_eng) != 0) {
		report("qbf_engine_create");
		g_eng = NULL;
		return NULL;
	}
	g_cells = cells;
	return g_eng;
}

typedef struct { void *p; size_t n; } dbuf_t;
static dbuf_t g_buf[8];

static void *dbuf(int i, size_t bytes)
{
	if (g_buf[i].n >= bytes) return g_buf[i].p;
	qbf_free(g_buf[i].p);
	g_buf[i].p = qbf_malloc(bytes);
	g_buf[i].n = g_buf[i].p ? bytes : 0;
	if (!g_buf[i].p) report("qbf_malloc");
	return g_buf[i].p;
}

/* ------------------------------------------------------------------ quadrature table cache */
/* Tables written by FourierI1..I4 are trusted (never re-hashed) until QuadI*_Delete. */
#define MAX_TRUSTED 64
static const void *g_trusted[MAX_TRUSTED];

static void trust(const void *p)
{
	for (int i = 0; i < MAX_TRUSTED; i++)
		if (!g_trusted[i]) { g_trusted[i] = p; return; }
}
static int is_trusted(const void *p)
{
	for (int i = 0; i < MAX_TRUSTED; i++)
		if (p && g_trusted[i] == p) return 1;
	return 0;
}
static void untrust(const void *p)
{
	for (int i = 0; i < MAX_TRUSTED; i++)
		if (g_trusted[i] == p) g_trusted[i] = NULL;
	for (int s = 0; s < NSLOT; s++)
		if (g_slot[s].p1 == p) g_slot[s].valid = 0;
}

static uint64_t hash_words(const void *a, size_t bytes, uint64_t h)
{
	const uint64_t *w = (const uint64_t *)a;
	for (size_t i = 0; i < bytes / 8; i++) h = (h ^ w[i]) * 0x100000001b3ULL + (h >> 29);
	return h;
}

static int need_upload(int slot, const void *p1, const void *p2, unsigned num, size_t b1, size_t b2)
{
	uint64_t h = 0;
	if (!is_trusted(p1)) h = hash_words(p2, b2, hash_words(p1, b1, 0xcbf29ce484222325ULL));
	slot_key_t *k = &g_slot[slot];
	if (k->valid && k->p1 == p1 && k->p2 == p2 && k->num == num && k->hash == h) return 0;
	k->p1 = p1; k->p2 = p2; k->num = num; k->hash = h; k->valid = 1;
	return 1;
}

static int use_i1(int slot, const quadI1_t *q)
{
	if (!need_upload(slot == QBF_TAB_I1 ? SLOT_I1 : SLOT_I4, q->psiR1, q->psiR2, q->num,
		sizeof(grid2_t) * q->num, sizeof(grid_t) * q->num)) return 0;
	if (qbf_set_i1_table(g_eng, slot, q->num, q->psiR1[0].data, q->psiR2[0].data)) {
		g_slot[slot == QBF_TAB_I1 ? SLOT_I1 : SLOT_I4].valid = 0;
		report("qbf_set_i1_table");
		return -1;
	}
	return 0;
}

static int use_i2(const quadI2_t *q)
{
	if (!need_upload(SLOT_I2, q->psiR1, q->psiR2, q->num, sizeof(grid_t) * q->num, sizeof(grid_t) * q->num)) return 0;
	if (qbf_set_i2_table(g_eng, q->num, q->psiR1[0].data, q->psiR2[0].data)) {
		g_slot[SLOT_I2].valid = 0;
		report("qbf_set_i2_table");
		return -1;
	}
	return 0;
}

static int use_i3(const quadI3_t *q)
{
	if (!need_upload(SLOT_I3, q->psiR1, q->psiR2, q->num, sizeof(cgrid_t) * q->num, sizeof(grid2_t) * q->num)) return 0;
	if (qbf_set_i3_table(g_eng, q->num, &q->psiR1[0].data[0][0], q->psiR2[0].data)) {
		g_slot[SLOT_I3].valid = 0;
		report("qbf_set_i3_table");
		return -1;
	}
	return 0;
}

static int use_all(const quadI1_t *q1, const quadI2_t *q2, const quadI3_t *q3, const quadI1_t *q4)
{
	return use_i1(QBF_TAB_I1, q1) || use_i2(q2) || use_i3(q3) || use_i1(QBF_TAB_I4, q4);
}

/* ------------------------------------------------------------------ quadrature.h */
void QuadI1_Allocate(const unsigned int num, quadI1_t *quad)
{
	quad->num = num;
	quad->psiR1 = fftw_malloc(num * sizeof(quad->psiR1[0]));
	quad->psiR2 = fftw_malloc(num * sizeof(quad->psiR2[0]));
}
void QuadI1_Delete(quadI1_t *quad)
{
	untrust(quad->psiR1);
	fftw_free(quad->psiR2);
	fftw_free(quad->psiR1);
	quad->num = 0;
}
void QuadI2_Allocate(const unsigned int num, quadI2_t *quad)
{
	quad->num = num;
	quad->psiR1 = fftw_malloc(num * sizeof(quad->psiR1[0]));
	quad->psiR2 = fftw_malloc(num * sizeof(quad->psiR2[0]));
}
void QuadI2_Delete(quadI2_t *quad)
{
	untrust(quad->psiR1);
	fftw_free(quad->psiR2);
	fftw_free(quad->psiR1);
	quad->num = 0;
}
void QuadI3_Allocate(const unsigned int num, quadI3_t *quad)
{
	quad->num = num;
	quad->psiR1 = fftw_malloc(num * sizeof(quad->psiR1[0]));
	quad->psiR2 = fftw_malloc(num * sizeof(quad->psiR2[0]));
}
void QuadI3_Delete(quadI3_t *quad)
{
	untrust(quad->psiR1);
	fftw_free(quad->psiR2);
	fftw_free(quad->psiR1);
	quad->num = 0;
}

void FourierI1(const unsigned int J, const double L, const double R, quadI1_t *quad)
{
	QuadI1_Allocate(J * J, quad);
	qbf_fourier_I1(N_GRID, J, L, R, quad->psiR1[0].data, quad->psiR2[0].data);
	trust(quad->psiR1);
}
void FourierI2(const unsigned int J, const double L, const double R, quadI2_t *quad)
{
	QuadI2_Allocate(J, quad);
	qbf_fourier_I2(N_GRID, J, L, R, quad->psiR1[0].data, quad->psiR2[0].data);
	trust(quad->psiR1);
}
void FourierI3(const unsigned int J, const double L, const double R, const unsigned int M, quadI3_t *quad)
{
	QuadI3_Allocate(J * M, quad);
	qbf_fourier_I3(N_GRID, J, L, R, M, &quad->psiR1[0].data[0][0], quad->psiR2[0].data);
	trust(quad->psiR1);
}
void FourierI4(const unsigned int J, const double L, const double R, quadI1_t *quad)
{
	QuadI1_Allocate(J, quad);
	qbf_fourier_I4(N_GRID, J, L, R, quad->psiR1[0].data, quad->psiR2[0].data);
	trust(quad->psiR1);
}
void LegendreGaussQuad(const unsigned int M, const double a, const double b, double *x, double *w)
{
	qbf_legendre_gauss(M, a, b, x, w);
}

/* ------------------------------------------------------------------ integrals.h */
void I1interm_Create(I1interm_t *interm) { interm->impl = NULL; }
void I1interm_Delete(I1interm_t *interm) { interm->impl = NULL; }
void I2interm_Create(I2interm_t *interm) { interm->impl = NULL; }
void I2interm_Delete(I2interm_t *interm) { interm->impl = NULL; }
void I3interm_Create(I3interm_t *interm) { interm->impl = NULL; }
void I3interm_Delete(I3interm_t *interm) { interm->impl = NULL; }

typedef int (*integral_fn)(qbf_engine *, const double *, const double *, const double *, double *);

static void run_integral(const char *name, int kind, int slot, const cgrid_t *f, const cgrid_t *g, const cgrid_t *h, cgrid_t *out)
{
	double *d = dbuf(0, 4 * sizeof(cgrid_t));
	if (!d) return;
	const size_t n = 2 * NN_;
	if (qbf_h2d(g_eng, d, f->data, sizeof(cgrid_t)) || qbf_h2d(g_eng, d + n, g->data, sizeof(cgrid_t)) ||
		qbf_h2d(g_eng, d + 2 * n, h->data, sizeof(cgrid_t))) { report(name); return; }
	int r;
	if (kind == 1) r = qbf_I1(g_eng, slot, d, d + n, d + 2 * n, d + 3 * n);
	else if (kind == 2) r = qbf_I2(g_eng, d, d + n, d + 2 * n, d + 3 * n);
	else r = qbf_I3(g_eng, d, d + n, d + 2 * n, d + 3 * n);
	if (r || qbf_d2h(g_eng, out->data, d + 3 * n, sizeof(cgrid_t))) report(name);
}

void I1integral(const cgrid_t *restrict_ f, const cgrid_t *restrict_ g, const cgrid_t *restrict_ h,
	const quadI1_t *restrict_ quad, I1interm_t *restrict_ interm, cgrid_t *restrict_ I1)
{
	if (!engine(1) || use_i1(QBF_TAB_I1, quad)) return;
	run_integral("I1integral", 1, QBF_TAB_I1, f, g, h, I1);
}

void I2integral(const cgrid_t *restrict_ f, const cgrid_t *restrict_ g, const cgrid_t *restrict_ h,
	const quadI2_t *restrict_ quad, I2interm_t *restrict_ interm, cgrid_t *restrict_ I2)
{
	if (!engine(1) || use_i2(quad)) return;
	run_integral("I2integral", 2, 0, f, g, h, I2);
}

void I3integral(const cgrid_t *restrict_ f, const cgrid_t *restrict_ g, const cgrid_t *restrict_ h,
	const quadI3_t *restrict_ quad, I3interm_t *restrict_ interm, cgrid_t *restrict_ I3)
{
	if (!engine(1) || use_i3(quad)) return;
	run_integral("I3integral", 3, 0, f, g, h, I3);
}

/* ------------------------------------------------------------------ collision.h */
void CdInterm_Create(CdInterm_t *interm) { interm->impl = NULL; }
void CdInterm_Delete(CdInterm_t *interm) { interm->impl = NULL; }
void CcInterm_Create(CcInterm_t *interm) { interm->impl = NULL; }
void CcInterm_Delete(CcInterm_t *interm) { interm->impl = NULL; }

static void run_operator(const char *name, int mode, const wignerF_t *W, wignerF_t *out)
{
	double *dW = dbuf(0, sizeof(wignerF_t)), *dC = dbuf(1, sizeof(wignerF_t));
	if (!dW || !dC) return;
	if (qbf_h2d(g_eng, dW, W, sizeof(wignerF_t)) || qbf_collision(g_eng, dW, 1, mode, dC, NULL, NULL, NULL, 1.0) ||
		qbf_d2h(g_eng, out, dC, sizeof(wignerF_t))) report(name);
}

void CdInt(const wignerF_t *restrict_ W, const quadI2_t *restrict_ quadI2, const quadI3_t *restrict_ quadI3,
	const quadI1_t *restrict_ quadI4, CdInterm_t *restrict_ interm, wignerF_t *restrict_ Cd)
{
	if (!engine(1) || use_i2(quadI2) || use_i3(quadI3) || use_i1(QBF_TAB_I4, quadI4)) return;
	run_operator("CdInt", QBF_CD, W, Cd);
}

void CcInt(const wignerF_t *restrict_ W, const quadI1_t *restrict_ quadI1, CcInterm_t *restrict_ interm,
	wignerF_t *restrict_ Cc)
{
	if (!engine(1) || use_i1(QBF_TAB_I1, quadI1)) return;
	run_operator("CcInt", QBF_CC, W, Cc);
}

/* ------------------------------------------------------------------ simulation.h */
void CollisionInterm_Create(collisionInterm_t *interm) { interm->impl = NULL; }
void CollisionInterm_Delete(collisionInterm_t *interm) { interm->impl = NULL; }

void CollisionStep(const wignerF_t *restrict_ W, const quadI1_t *restrict_ quadI1, const quadI2_t *restrict_ quadI2,
	const quadI3_t *restrict_ quadI3, const quadI1_t *restrict_ quadI4,
	const double dt, const double Bext[3], collisionInterm_t *restrict_ interm, wignerF_t *restrict_ Wnext)
{
	if (!engine(1) || use_all(quadI1, quadI2, quadI3, quadI4)) return;
	double *dW = dbuf(0, sizeof(wignerF_t)), *dN = dbuf(1, sizeof(wignerF_t)), *dB = dbuf(2, 3 * sizeof(double));
	if (!dW || !dN || !dB) return;
	if (qbf_h2d(g_eng, dW, W, sizeof(wignerF_t)) || qbf_h2d(g_eng, dB, Bext, 3 * sizeof(double)) ||
		qbf_collision_step(g_eng, dW, 1, dt, dB, dN) || qbf_d2h(g_eng, Wnext, dN, sizeof(wignerF_t)))
		report("CollisionStep");
}

typedef struct {
	double vgrid[N_GRID];
} inhom_impl_t;

void InhomStepInterm_Create(const unsigned int numVol, const double L, inhomStepInterm_t *interm)
{
	interm->numVol = numVol;
	interm->L = L;
	inhom_impl_t *im = calloc(1, sizeof(inhom_impl_t));
	const double scale = (2 * L) / N_GRID;
	for (int k = 0; k < N_GRID; k++) im->vgrid[k] = k < N_GRID / 2 ? scale * k : scale * (k - N_GRID);
	interm->impl = im;
	engine((int)numVol);
}

void InhomStepInterm_Delete(inhomStepInterm_t *interm)
{
	free(interm->impl);
	interm->impl = NULL;
}

void InhomStepPeriodic(const wignerV_t *restrict_ W, const quadI1_t *restrict_ quadI1, const quadI2_t *restrict_ quadI2,
	const quadI3_t *restrict_ quadI3, const quadI1_t *restrict_ quadI4,
	const double h, const double dt, const double *restrict_ Bext, inhomStepInterm_t *restrict_ interm, wignerV_t *restrict_ Wnext)
{
	const unsigned nv = interm->numVol;
	if (!engine((int)nv) || use_all(quadI1, quadI2, quadI3, quadI4)) return;
	const size_t sz = nv * sizeof(wignerV_t);
	double *dW = dbuf(0, sz), *dN = dbuf(1, sz), *dB = dbuf(2, 3 * nv * sizeof(double));
	if (!dW || !dN || !dB) return;
	if (qbf_h2d(g_eng, dW, W, sz) || qbf_h2d(g_eng, dB, Bext, 3 * nv * sizeof(double)) ||
		qbf_inhom_step_periodic(g_eng, dW, (int)nv, interm->L, h, dt, dB, dN) || qbf_d2h(g_eng, Wnext, dN, sz))
		report("InhomStepPeriodic");
}

void InhomStepDirichlet(const wignerV_t *restrict_ W, const quadI1_t *restrict_ quadI1, const quadI2_t *restrict_ quadI2,
	const quadI3_t *restrict_ quadI3, const quadI1_t *restrict_ quadI4,
	const double h, const double dt, const double *restrict_ Bext, inhomStepInterm_t *restrict_ interm, wignerV_t *restrict_ Wnext)
{
	/* half transport with Dirichlet ends, collision on the interior cells only (the end cells
	 * keep their transported values, src/simulation.c:715), half transport (src/simulation.c:562-830),
	 * all on the GPU */
	const unsigned nv = interm->numVol;
	if (!engine((int)nv) || use_all(quadI1, quadI2, quadI3, quadI4)) return;
	const size_t sz = nv * sizeof(wignerV_t);
	double *dW = dbuf(0, sz), *dN = dbuf(1, sz), *dB = dbuf(2, 3 * nv * sizeof(double));
	if (!dW || !dN || !dB) return;
	if (qbf_h2d(g_eng, dW, W, sz) || qbf_h2d(g_eng, dB, Bext, 3 * nv * sizeof(double)) ||
		qbf_inhom_step_dirichlet(g_eng, dW, (int)nv, interm->L, h, dt, dB, dN) || qbf_d2h(g_eng, Wnext, dN, sz))
		report("InhomStepDirichlet");
}

void InhomStepMaxwell(const wignerV_t *restrict_ W, const quadI1_t *restrict_ quadI1, const quadI2_t *restrict_ quadI2,
	const quadI3_t *restrict_ quadI3, const quadI1_t *restrict_ quadI4,
	const double h, const double dt, const double lambda, const wignerV_t *restrict_ WMaxwL, const wignerV_t *restrict_ WMaxwR,
	const double *restrict_ Bext, inhomStepInterm_t *restrict_ interm, wignerV_t *restrict_ Wnext)
{
	/* half transport with Maxwell reflection, collision of every cell, half transport
	 * (src/simulation.c:882-1165), all on the GPU; the boundary states travel as given and are
	 * moved to the Maxwell velocity grid and normalised on the device */
	const unsigned nv = interm->numVol;
	if (!engine((int)nv) || use_all(quadI1, quadI2, quadI3, quadI4)) return;
	const size_t sz = nv * sizeof(wignerV_t);
	double *dW = dbuf(0, sz), *dN = dbuf(1, sz), *dB = dbuf(2, 3 * nv * sizeof(double));
	double *dM = dbuf(6, 2 * sizeof(wignerV_t));
	if (!dW || !dN || !dB || !dM) return;
	double *dR = dM + sizeof(wignerV_t) / sizeof(double);
	if (qbf_h2d(g_eng, dW, W, sz) || qbf_h2d(g_eng, dB, Bext, 3 * nv * sizeof(double)) ||
		qbf_h2d(g_eng, dM, WMaxwL, sizeof(wignerV_t)) || qbf_h2d(g_eng, dR, WMaxwR, sizeof(wignerV_t)) ||
		qbf_inhom_step_maxwell(g_eng, dW, (int)nv, interm->L, h, dt, lambda, dM, dR, dB, dN) ||
		qbf_d2h(g_eng, Wnext, dN, sz))
		report("InhomStepMaxwell");
}

void SimulationHomInterm_Create(simulationHomInterm_t *interm) { interm->impl = NULL; }
void SimulationHomInterm_Delete(simulationHomInterm_t *interm) { interm->impl = NULL; }

void SimulationHomogeneous(const wignerV_t *restrict_ W0, const quadI1_t *restrict_ quadI1, const quadI2_t *restrict_ quadI2,
	const quadI3_t *restrict_ quadI3, const quadI1_t *restrict_ quadI4,
	const double dt, const unsigned int numsteps, const double Bext[3], simulationHomInterm_t *restrict_ interm, wignerV_t *restrict_ Wevolv)
{
	memcpy(Wevolv, W0, sizeof(wignerV_t));
	if (numsteps < 2) return;
	if (!engine(1) || use_all(quadI1, quadI2, quadI3, quadI4)) return;
	double *dP = dbuf(0, sizeof(wignerV_t)), *dF = dbuf(3, sizeof(wignerF_t)), *dF2 = dbuf(4, sizeof(wignerF_t));
	double *dB = dbuf(2, 3 * sizeof(double));
	if (!dP || !dF || !dF2 || !dB) return;
	if (qbf_h2d(g_eng, dP, W0, sizeof(wignerV_t)) || qbf_h2d(g_eng, dB, Bext, 3 * sizeof(double))) {
		report("SimulationHomogeneous");
		return;
	}
	/* the state stays on the GPU; every frame is copied out (src/simulation.c:1219-1281) */
	for (unsigned it = 1; it < numsteps; it++) {
		if (qbf_phys_to_fourier(g_eng, dP, 4, dF) || qbf_collision_step(g_eng, dF, 1, dt, dB, dF2) ||
			qbf_fourier_to_phys(g_eng, dF2, 4, dP) || qbf_d2h(g_eng, &Wevolv[it], dP, sizeof(wignerV_t))) {
			report("SimulationHomogeneous");
			return;
		}
	}
}

void SimulationInhomogeneousPeriodic(const wignerV_t *restrict_ W0, const quadI1_t *restrict_ quadI1, const quadI2_t *restrict_ quadI2,
	const quadI3_t *restrict_ quadI3, const quadI1_t *restrict_ quadI4,
	const double h, const double dt, const unsigned int numsteps, const double *restrict_ Bext, inhomStepInterm_t *restrict_ interm, wignerV_t *restrict_ Wevolv)
{
	const unsigned nv = interm->numVol;
	memcpy(Wevolv, W0, nv * sizeof(wignerV_t));
	if (numsteps < 2) return;
	if (!engine((int)nv) || use_all(quadI1, quadI2, quadI3, quadI4)) return;
	const size_t sz = nv * sizeof(wignerV_t);
	double *dA = dbuf(5, sz), *dB = dbuf(2, 3 * nv * sizeof(double));
	if (!dA || !dB) return;
	/* the state stays on the GPU; every frame is copied out (src/simulation.c:1288-1301) */
	if (qbf_h2d(g_eng, dA, W0, sz) || qbf_h2d(g_eng, dB, Bext, 3 * nv * sizeof(double)) ||
		qbf_simulate_periodic(g_eng, dA, (int)nv, interm->L, h, dt, numsteps, 1, dB, (double *)Wevolv))
		report("SimulationInhomogeneousPeriodic");
}

void SimulationInhomogeneousDirichlet(const wignerV_t *restrict_ W0, const quadI1_t *restrict_ quadI1, const quadI2_t *restrict_ quadI2,
	const quadI3_t *restrict_ quadI3, const quadI1_t *restrict_ quadI4,
	const double h, const double dt, const unsigned int numsteps, const double *restrict_ Bext, inhomStepInterm_t *restrict_ interm, wignerV_t *restrict_ Wevolv)
{
	const unsigned nv = interm->numVol;
	memcpy(Wevolv, W0, nv * sizeof(wignerV_t));
	if (numsteps < 2) return;
	if (!engine((int)nv) || use_all(quadI1, quadI2, quadI3, quadI4)) return;
	const size_t sz = nv * sizeof(wignerV_t);
	double *dA = dbuf(5, sz), *dB = dbuf(2, 3 * nv * sizeof(double));
	if (!dA || !dB) return;
	/* the state stays on the GPU; every frame is copied out (src/simulation.c:1308-1320) */
	if (qbf_h2d(g_eng, dA, W0, sz) || qbf_h2d(g_eng, dB, Bext, 3 * nv * sizeof(double)) ||
		qbf_simulate_dirichlet(g_eng, dA, (int)nv, interm->L, h, dt, numsteps, 1, dB, (double *)Wevolv))
		report("SimulationInhomogeneousDirichlet");
}

void SimulationInhomogeneousMaxwell(const wignerV_t *restrict_ W0, const quadI1_t *restrict_ quadI1, const quadI2_t *restrict_ quadI2,
	const quadI3_t *restrict_ quadI3, const quadI1_t *restrict_ quadI4,
	const double h, const double dt, const double lambda, const wignerV_t *restrict_ WMaxwL, const wignerV_t *restrict_ WMaxwR,
	const unsigned int numsteps, const double *restrict_ Bext, inhomStepInterm_t *restrict_ interm, wignerV_t *restrict_ Wevolv)
{
	const unsigned nv = interm->numVol;
	memcpy(Wevolv, W0, nv * sizeof(wignerV_t));
	if (numsteps < 2) return;
	if (!engine((int)nv) || use_all(quadI1, quadI2, quadI3, quadI4)) return;
	const size_t sz = nv * sizeof(wignerV_t);
	double *dA = dbuf(5, sz), *dB = dbuf(2, 3 * nv * sizeof(double)), *dM = dbuf(6, 2 * sizeof(wignerV_t));
	if (!dA || !dB || !dM) return;
	double *dR = dM + sizeof(wignerV_t) / sizeof(double);
	/* the state stays on the GPU; every frame is copied out (src/simulation.c:1327-1341) */
	if (qbf_h2d(g_eng, dA, W0, sz) || qbf_h2d(g_eng, dB, Bext, 3 * nv * sizeof(double)) ||
		qbf_h2d(g_eng, dM, WMaxwL, sizeof(wignerV_t)) || qbf_h2d(g_eng, dR, WMaxwR, sizeof(wignerV_t)) ||
		qbf_simulate_maxwell(g_eng, dA, (int)nv, interm->L, h, dt, lambda, dM, dR, numsteps, 1, dB, (double *)Wevolv))
		report("SimulationInhomogeneousMaxwell");
}

/* ------------------------------------------------------------------ finite_volume.h
 * Single-line host entry points (src/finite_volume.c).  The periodic one runs on the GPU
 * through the batched transport kernel; the single-line Dirichlet and Maxwell variants are host
 * utilities (the Dirichlet and Maxwell steps above run their batched kernels on the GPU). */
static double minmod(double a, double b)
{
	return 0.5 * (copysign(1, a) + copysign(1, b)) * fmin(fabs(a), fabs(b));
}

void SlopeLimiterStepPeriodic(const double h, const double dt, const double A, const unsigned int N, const double *restrict_ Un, double *restrict_ Un1)
{
	if (!engine(1)) return;
	double *d = dbuf(7, 2 * N * sizeof(double));
	if (!d) return;
	if (qbf_h2d(g_eng, d, Un, N * sizeof(double)) || qbf_transport_line_periodic(g_eng, d, (int)N, A, h, dt, d + N) ||
		qbf_d2h(g_eng, Un1, d + N, N * sizeof(double))) report("SlopeLimiterStepPeriodic");
}

void SlopeLimiterStepDirichlet(const double h, const double dt, const double A, const unsigned int N, const double *restrict_ Un, double *restrict_ Un1)
{
	const double Aabs = fabs(A), nu = dt * A / h;
	const double S = 0.5 * A * (copysign(1, nu) - nu), Sabs = 0.5 * Aabs * (copysign(1, nu) - nu);
	const double scale = dt / (2 * h);
	/* h*beta_j = minmod(alpha_j, alpha_{j-1}), alpha_{N-1} = 0, h*beta at both ends = 0 */
	double *hb = malloc(N * sizeof(double));
	hb[0] = 0;
	hb[N - 1] = 0;
	for (unsigned j = 1; j + 1 < N; j++) hb[j] = minmod(Un[j + 1] - Un[j], Un[j] - Un[j - 1]);
	for (unsigned j = 1; j + 1 < N; j++)
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
	free(hb);
}

double CalculateRightwardFlux(const unsigned int m, const double *restrict_ A, const double *restrict_ U)
{
	double flux = 0;
	for (unsigned i = 0; i < m / 2; i++) flux += A[i] * U[i];
	return flux;
}

double CalculateLeftwardFlux(const unsigned int m, const double *restrict_ A, const double *restrict_ U)
{
	double flux = 0;
	for (unsigned i = m / 2; i < m; i++) flux += fabs(A[i]) * U[i];
	return flux;
}

void SlopeLimiterStepMaxwell(const double h, const double dt, const unsigned int m, const double *restrict_ A, const unsigned int N,
	const double lambda, const double fluxL, const double fluxR, const double *restrict_ UmaxwL, const double *restrict_ UmaxwR,
	const double *restrict_ Un, double *restrict_ Un1)
{
	/* ghost cells: two copies of the Maxwell-reflected state on each side (src/finite_volume.c:247-292) */
	double *V = malloc((N + 4) * sizeof(double)), *hb = malloc((N + 2) * sizeof(double));
	const double scale = dt / (2 * h);
	for (unsigned i = 0; i < m; i++) {
		const int out_right = i < m / 2;     /* non-negative velocities: outgoing at the right end */
		const double gl = out_right ? (1 - lambda) * Un[m - 1 - i] + lambda * fluxL * UmaxwL[i] : Un[i];
		const double gr = out_right ? Un[i + m * (N - 1)] : (1 - lambda) * Un[m - 1 - i + m * (N - 1)] + lambda * fluxR * UmaxwR[i];
		V[0] = V[1] = gl;
		for (unsigned j = 0; j < N; j++) V[j + 2] = Un[i + m * j];
		V[N + 2] = V[N + 3] = gr;
		for (unsigned j = 0; j < N + 2; j++) hb[j] = minmod(V[j + 2] - V[j + 1], V[j + 1] - V[j]);   /* cell j-1 */
		const double Aabs = fabs(A[i]), nu = dt * A[i] / h;
		const double S = 0.5 * A[i] * (copysign(1, nu) - nu), Sabs = 0.5 * Aabs * (copysign(1, nu) - nu);
		for (unsigned j = 0; j < N; j++) {
			const double vm = V[j + 1], v0 = V[j + 2], vp = V[j + 3];
			const double bm = hb[j], b0 = hb[j + 1], bp = hb[j + 2];
			Un1[i + m * j] = v0 + scale * (
				-A[i] * (vp - vm) + Aabs * (vp - 2 * v0 + vm) - S * (bp - bm) + Sabs * (bp - 2 * b0 + bm));
		}
	}
	free(hb); free(V);
}

/* ------------------------------------------------------------------ grid.h (host grid utilities) */
void MultiplyPointwiseRR(const grid_t *restrict_ f, const grid_t *restrict_ g, grid_t *restrict_ ret)
{
	for (int i = 0; i < NN_; i++) ret->data[i] = f->data[i] * g->data[i];
}
void MultiplyPointwiseRC(const grid_t *restrict_ f, const cgrid_t *restrict_ g, cgrid_t *restrict_ ret)
{
	for (int i = 0; i < NN_; i++) { ret->data[i][0] = f->data[i] * g->data[i][0]; ret->data[i][1] = f->data[i] * g->data[i][1]; }
}
void MultiplyPointwiseCC(const cgrid_t *restrict_ f, const cgrid_t *restrict_ g, cgrid_t *restrict_ ret)
{
	for (int i = 0; i < NN_; i++) {
		ret->data[i][0] = f->data[i][0] * g->data[i][0] - f->data[i][1] * g->data[i][1];
		ret->data[i][1] = f->data[i][0] * g->data[i][1] + f->data[i][1] * g->data[i][0];
	}
}
void MultiplyPointwiseRR2(const grid2_t *restrict_ f, const grid2_t *restrict_ g, grid2_t *restrict_ ret)
{
	for (int i = 0; i < PP_; i++) ret->data[i] = f->data[i] * g->data[i];
}
void MultiplyPointwiseCC2(const cgrid2_t *restrict_ f, const cgrid2_t *restrict_ g, cgrid2_t *restrict_ ret)
{
	for (int i = 0; i < PP_; i++) {
		ret->data[i][0] = f->data[i][0] * g->data[i][0] - f->data[i][1] * g->data[i][1];
		ret->data[i][1] = f->data[i][0] * g->data[i][1] + f->data[i][1] * g->data[i][0];
	}
}
void MultiplyPointwiseAssignRR(const grid_t *restrict_ f, grid_t *restrict_ ret)
{
	for (int i = 0; i < NN_; i++) ret->data[i] *= f->data[i];
}
void MultiplyPointwiseAssignRC(const grid_t *restrict_ f, cgrid_t *restrict_ ret)
{
	for (int i = 0; i < NN_; i++) { ret->data[i][0] *= f->data[i]; ret->data[i][1] *= f->data[i]; }
}
void MultiplyPointwiseAssignRC2(const grid2_t *restrict_ f, cgrid2_t *restrict_ ret)
{
	for (int i = 0; i < PP_; i++) { ret->data[i][0] *= f->data[i]; ret->data[i][1] *= f->data[i]; }
}
void MultiplyPointwiseAssignCC2(const cgrid2_t *restrict_ f, cgrid2_t *restrict_ ret)
{
	for (int i = 0; i < PP_; i++) {
		const double re = ret->data[i][0], im = ret->data[i][1];
		ret->data[i][0] = f->data[i][0] * re - f->data[i][1] * im;
		ret->data[i][1] = f->data[i][0] * im + f->data[i][1] * re;
	}
}
/* frequency block [0..N/2-1, -N/2..-1]^2 <-> [0..N-1, -N..-1]^2 (src/grid.c:171-210) */
void GridZeroPad(const cgrid_t *restrict_ f, cgrid2_t *restrict_ ret)
{
	memset(ret->data, 0, sizeof(ret->data));
	for (int l = 0; l < N_GRID; l++) {
		const int lp = l < N_GRID / 2 ? l : l + N_GRID;
		memcpy(&ret->data[2 * N_GRID * lp], &f->data[N_GRID * l], (N_GRID / 2) * sizeof(fftw_complex));
		memcpy(&ret->data[3 * N_GRID / 2 + 2 * N_GRID * lp], &f->data[N_GRID / 2 + N_GRID * l], (N_GRID / 2) * sizeof(fftw_complex));
	}
}
void GridExtract(const cgrid2_t *restrict_ f, cgrid_t *restrict_ ret)
{
	for (int l = 0; l < N_GRID; l++) {
		const int lp = l < N_GRID / 2 ? l : l + N_GRID;
		memcpy(&ret->data[N_GRID * l], &f->data[2 * N_GRID * lp], (N_GRID / 2) * sizeof(fftw_complex));
		memcpy(&ret->data[N_GRID / 2 + N_GRID * l], &f->data[3 * N_GRID / 2 + 2 * N_GRID * lp], (N_GRID / 2) * sizeof(fftw_complex));
	}
}

/* ------------------------------------------------------------------ util.h */
int ReadData(const char *filename, void *data, const size_t size, const size_t n)
{
	FILE *fd = fopen(filename, "rb");
	if (!fd) {
		fprintf(stderr, "ReadData: cannot open '%s'\n", filename);
		return -1;
	}
	fseek(fd, 0, SEEK_END);
	const long filesize = ftell(fd);
	rewind(fd);
	if (filesize < 0 || (size_t)filesize != n * size) {
		fprintf(stderr, "ReadData: '%s' holds %ld bytes, %zu expected\n", filename, filesize, n * size);
		fclose(fd);
		return -2;
	}
	if (fread(data, size, n, fd) != n) {
		fprintf(stderr, "ReadData: short read from '%s'\n", filename);
		fclose(fd);
		return -3;
	}
	fclose(fd);
	return 0;
}

int WriteData(const char *filename, const void *data, const size_t size, const size_t n, const bool append)
{
	FILE *fd = fopen(filename, append ? "ab" : "wb");
	if (!fd) {
		fprintf(stderr, "WriteData: cannot open '%s'\n", filename);
		return -1;
	}
	if (fwrite(data, size, n, fd) != n) {
		fprintf(stderr, "WriteData: short write to '%s'\n", filename);
		fclose(fd);
		return -3;
	}
	fclose(fd);
	return 0;
}

/* ------------------------------------------------------------------ fftw3.h subset */
struct qbf_fftw_plan_s { int n0, n1; double *in; fftw_complex *out; };

void *fftw_malloc(size_t n)
{
	void *p = NULL;
	if (posix_memalign(&p, 64, n ? n : 64)) return NULL;
	return p;
}
void fftw_free(void *p) { free(p); }
void fftw_cleanup(void) {}

fftw_plan fftw_plan_dft_r2c_2d(int n0, int n1, double *in, fftw_complex *out, unsigned flags)
{
	if (n0 != N_GRID || n1 != N_GRID) {
		fprintf(stderr, "libqbf: fftw_plan_dft_r2c_2d supports %d x %d only\n", N_GRID, N_GRID);
		return NULL;
	}
	struct qbf_fftw_plan_s *p = malloc(sizeof(*p));
	p->n0 = n0; p->n1 = n1; p->in = in; p->out = out;
	return p;
}

void fftw_execute(const fftw_plan p)
{
	/* forward DFT of the real grid on the GPU (scaled by 1/N^2 there; the power-of-two
	 * rescale below is exact), half-spectrum layout out[l][0..N/2] */
	if (!p || !engine(1)) return;
	double *dP = dbuf(0, sizeof(grid_t)), *dF = dbuf(3, sizeof(cgrid_t));
	cgrid_t *full = malloc(sizeof(cgrid_t));
	if (dP && dF && !qbf_h2d(g_eng, dP, p->in, sizeof(grid_t)) && !qbf_phys_to_fourier(g_eng, dP, 1, dF) &&
		!qbf_d2h(g_eng, full, dF, sizeof(cgrid_t))) {
		const double s = (double)N_GRID * N_GRID;
		for (int l = 0; l < N_GRID; l++)
			for (int k = 0; k <= N_GRID / 2; k++) {
				p->out[(N_GRID / 2 + 1) * l + k][0] = full->data[N_GRID * l + k][0] * s;
				p->out[(N_GRID / 2 + 1) * l + k][1] = full->data[N_GRID * l + k][1] * s;
			}
	} else {
		report("fftw_execute");
	}
	free(full);
}

void fftw_destroy_plan(fftw_plan p) { free(p); }
