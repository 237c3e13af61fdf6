/*
 * qbf.h -- device-resident C API of the MI355X collision-plus-transport engine.
 *
 * This is the native layer under the drop-in reference headers (collision.h,
 * integrals.h, simulation.h, finite_volume.h of cmendl/inhom-quantum-boltzmann-fourier-c,
 * re-declared in this directory).  It has no counterpart in the reference: it
 * exposes batched entry points over device memory so that a time loop keeps
 * the state in HBM (the reference's entry points take host memory and one cell
 * at a time).
 *
 * Layouts (all device pointers, double precision, complex = interleaved pair):
 *   Fourier state  [ncell][4][N][N] complex   (wignerF_t per cell, include/wigner.h:47-50)
 *   physical state [ncell][4][N][N] real      (wignerV_t per cell, include/wigner.h:35-41)
 *   index within a grid: k + N*l, l = slow index (v_x), k = fast index (v_y).
 * Every function returns 0 on success and a negative code on error; the message
 * is available from qbf_last_error().  Calls are asynchronous on the engine's
 * stream; qbf_sync() waits for completion.
 */
#ifndef QBF_H
#define QBF_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qbf_engine qbf_engine;

/* arithmetic of the convolution kernels.  QBF_BF16X3 (N = 64 only): as QBF_FP32, except that the
 * line DFTs of the I3 K-operator (the dominant work) are GEMMs on the matrix cores with BF16 hi/lo
 * split operands (hi*hi + hi*lo + lo*hi, FP32 accumulation); no counterpart in the reference. */
enum { QBF_FP64 = 0, QBF_FP32 = 1, QBF_BF16X3 = 2 };
enum { QBF_CC = 1, QBF_CD = 2, QBF_C = 3 };        /* which collision operator */
enum { QBF_TAB_I1 = 0, QBF_TAB_I4 = 1 };           /* I1-type table slots */

/* Quadrature tables in the reference's host layout (include/quadrature.h:35-77), P = 2N:
 *   I1/I4: psiR1[num][P*P] real, psiR2[num][N*N] real
 *   I2:    psiR1[num][N*N] real, psiR2[num][N*N] real
 *   I3:    psiR1[num][N*N] complex, psiR2[num][P*P] real */
typedef struct {
	unsigned numI1; const double *I1_psiR1, *I1_psiR2;
	unsigned numI2; const double *I2_psiR1, *I2_psiR2;
	unsigned numI3; const double *I3_psiR1, *I3_psiR2;
	unsigned numI4; const double *I4_psiR1, *I4_psiR2;
} qbf_quad_tables;

const char *qbf_last_error(void);

/* N = 16, 32, 64 (padded grids held in registers/LDS by one workgroup) or 128 (padded 256 x 256 grids
 * in global memory).  The reference fixes N at compile time (include/grid.h:34). */
int  qbf_engine_create(int N, int precision, int maxCells, qbf_engine **out);
void qbf_engine_destroy(qbf_engine *e);
int  qbf_engine_N(const qbf_engine *e);
int  qbf_engine_precision(const qbf_engine *e);
/* run on an existing stream (e.g. torch.cuda.current_stream().cuda_stream); NULL = own stream */
int  qbf_set_stream(qbf_engine *e, void *hip_stream);
int  qbf_sync(qbf_engine *e);
void *qbf_get_stream(qbf_engine *e);     /* the stream the engine launches on */

/* upload all four tables (host memory) */
int qbf_set_quadrature(qbf_engine *e, const qbf_quad_tables *q);
/* upload a single table set (host memory) */
int qbf_set_i1_table(qbf_engine *e, int slot, unsigned num, const double *psiR1, const double *psiR2);
int qbf_set_i2_table(qbf_engine *e, unsigned num, const double *psiR1, const double *psiR2);
int qbf_set_i3_table(qbf_engine *e, unsigned num, const double *psiR1, const double *psiR2);

/* out = base + cf * ( C_mode(W) + 2 (Bext x w)(Wcb) )   for ncell Fourier states.
 * base, Wcb may be NULL (then 0 / no magnetic term); Bext is [ncell][3] (device). */
int qbf_collision(qbf_engine *e, const double *dW, int ncell, int mode, double *dOut,
	const double *dBase, const double *dWcb, const double *dBext, double cf);

/* explicit midpoint collision step on ncell Fourier states (src/simulation.c:105-136) */
int qbf_collision_step(qbf_engine *e, const double *dW, int ncell, double dt, const double *dBext, double *dWnext);

/* single integrals on one N x N complex grid each (src/integrals.c) */
int qbf_I1(qbf_engine *e, int slot, const double *df, const double *dg, const double *dh, double *dOut);
int qbf_I2(qbf_engine *e, const double *df, const double *dg, const double *dh, double *dOut);
int qbf_I3(qbf_engine *e, const double *df, const double *dg, const double *dh, double *dOut);

/* physical real N x N grids <-> Fourier (scaled by 1/N^2; inverse: unnormalised, real part) */
int qbf_phys_to_fourier(qbf_engine *e, const double *dPhys, int ngrid, double *dFourier);
int qbf_fourier_to_phys(qbf_engine *e, const double *dFourier, int ngrid, double *dPhys);

/* slope-limited transport, periodic in x, over ncell physical states (src/finite_volume.c:52-104) */
int qbf_transport_periodic(qbf_engine *e, const double *dW, int ncell, double L, double h, double dt, double *dWout);

/* slope-limited transport with Dirichlet ends over ncell >= 2 physical states (src/finite_volume.c:110-170) */
int qbf_transport_dirichlet(qbf_engine *e, const double *dW, int ncell, double L, double h, double dt, double *dWout);

/* one Strang step with Dirichlet ends: half transport, collision of the interior cells
 * 1 .. ncell-2 (the end cells keep their transported values), half transport (src/simulation.c:562-830) */
int qbf_inhom_step_dirichlet(qbf_engine *e, const double *dW, int ncell, double L, double h, double dt,
	const double *dBext, double *dWnext);

/* slope-limited transport with Maxwell reflection at both ends (src/finite_volume.c:241-323 on the
 * Maxwell velocity grid, src/simulation.c:837-1051): the zero x-velocity row is kept, the incoming
 * ghost states are (1 - lambda) * reflected end cell + lambda * outgoing trace flux * boundary state
 * normalised by its incoming trace flux.  dWMaxwL, dWMaxwR: one physical state each, standard grid. */
int qbf_transport_maxwell(qbf_engine *e, const double *dW, int ncell, double L, double h, double dt, double lambda,
	const double *dWMaxwL, const double *dWMaxwR, double *dWout);

/* one Strang step with Maxwell boundaries: half transport, collision of every cell, half transport
 * (replaces InhomStepMaxwell's compute, src/simulation.c:882-1165) */
int qbf_inhom_step_maxwell(qbf_engine *e, const double *dW, int ncell, double L, double h, double dt, double lambda,
	const double *dWMaxwL, const double *dWMaxwR, const double *dBext, double *dWnext);

/* slope-limited transport of one periodic line of n cells with velocity A */
int qbf_transport_line_periodic(qbf_engine *e, const double *dU, int n, double A, double h, double dt, double *dOut);

/* one Strang-split periodic step: half transport, collision, half transport (src/simulation.c:307-555) */
int qbf_inhom_step_periodic(qbf_engine *e, const double *dW, int ncell, double L, double h, double dt,
	const double *dBext, double *dWnext);

/* After qbf_inhom_step_periodic on ncell cells: the state its collision step started from (after
 * the first half transport; dWin may be NULL) and the collision operator at that state, C(W) plus
 * the magnetic term, physical [ncell][4][N][N], taken from the step's FP64 midpoint state as
 * 2 (W_mid - W) / dt (no further evaluation).  For anchoring a surrogate between exact steps;
 * no counterpart in the reference. */
int qbf_last_collision(qbf_engine *e, int ncell, double *dWin, double *dC);

/* One periodic Strang step (qbf_inhom_step_periodic with these device buffers) captured as a HIP
 * graph, for callers that replay a fixed step without per-launch host work; no counterpart in the
 * reference.  The step runs once eagerly first (dOut holds its result), then is captured on the
 * engine's stream, which must not be the null stream (qbf_set_stream).  Needs a fixed paired-path
 * choice (qbf_set_real_pairs mode 0 or 2) and no buffer growth during the capture.  Launches replay
 * on the capture stream, asynchronously, ordered after the work queued on the engine's current
 * stream and before the work queued there afterwards.  The graph uses the engine's buffers: while
 * it is alive, calls that would grow a buffer or drop a cached table fail (-1); after
 * qbf_engine_destroy the graph is dead (qbf_graph_launch fails, qbf_graph_destroy frees it). */
typedef struct qbf_graph qbf_graph;
int  qbf_graph_capture_inhom_step_periodic(qbf_engine *e, const double *dW, int ncell, double L, double h,
	double dt, const double *dBext, double *dOut, qbf_graph **g);
int  qbf_graph_launch(qbf_graph *g);
void qbf_graph_destroy(qbf_graph *g);
/* The same for any sequence of calls on the engine's stream: fn(ctx) runs once eagerly, then once
 * under capture (it must return 0; its calls may not allocate, copy between host and device or
 * synchronise while captured; the engine's own entry points return -1 where they would). */
int  qbf_graph_capture_calls(qbf_engine *e, int (*fn)(void *ctx), void *ctx, qbf_graph **g);

/* numsteps frames of the periodic problem from dW0 (device), the state kept on the device
 * (src/simulation.c:1288-1301): frame 0 = W0, frame it = qbf_inhom_step_periodic of frame it-1
 * (with the engine's conservation setting).  Frames 0, stride, 2 stride, ... < numsteps are
 * copied to hostFrames (host, (numsteps - 1) / stride + 1 frames of ncell cells); the call
 * returns when they are there.  stride = 1 gives the reference's Wevolv. */
int qbf_simulate_periodic(qbf_engine *e, const double *dW0, int ncell, double L, double h, double dt,
	unsigned numsteps, unsigned stride, const double *dBext, double *hostFrames);

/* The same loop with Dirichlet ends (frame it = qbf_inhom_step_dirichlet of frame it-1) and with
 * Maxwell boundaries (qbf_inhom_step_maxwell; dWMaxwL, dWMaxwR as there), replacing the per-step
 * host round trips of SimulationInhomogeneousDirichlet / Maxwell (src/simulation.c:1308-1341). */
int qbf_simulate_dirichlet(qbf_engine *e, const double *dW0, int ncell, double L, double h, double dt,
	unsigned numsteps, unsigned stride, const double *dBext, double *hostFrames);
int qbf_simulate_maxwell(qbf_engine *e, const double *dW0, int ncell, double L, double h, double dt, double lambda,
	const double *dWMaxwL, const double *dWMaxwR, unsigned numsteps, unsigned stride, const double *dBext,
	double *hostFrames);

/* qbf_simulate_periodic without any device allocation; no counterpart in the reference.  For callers
 * that cannot allocate (tight stepping loops, code that must not call hipMalloc).  The two state
 * buffers of the loop come from a static device array of 2 * QBF_NOALLOC_STATE_BYTES (64 MiB) that
 * each device holds from the library's first use on it until the process ends, so a state may hold
 * at most QBF_NOALLOC_STATE_BYTES (ncell * 4 * N * N * 8 bytes: 1024 cells at N = 32, 256 at
 * N = 64, 64 at N = 128; larger states return -1).  The engine's own buffers and combination
 * tables must already be sized for ncell cells (one qbf_inhom_step_periodic or
 * qbf_simulate_periodic call on that many cells does it); otherwise the call returns -1 instead of
 * allocating.  Calls on the engine's device (the current device) share the array: calls that
 * overlap on one device race.  Results equal qbf_simulate_periodic's bit for bit. */
#define QBF_NOALLOC_STATE_BYTES (32u << 20)
int qbf_simulate_periodic_noalloc(qbf_engine *e, const double *dW0, int ncell, double L, double h, double dt,
	unsigned numsteps, unsigned stride, const double *dBext, double *hostFrames);

/* FP64 conservation correction of a collision step (physical states, [ncell][4][N][N]):
 * Wnew is replaced by Wold + P(Wnew - Wold), where P removes, per cell, the selected moments
 * of the increment with a |w0|-weighted least-squares correction.  mask: QBF_CONS_MASS (mass
 * and the three spin densities), QBF_CONS_MOMENTUM, QBF_CONS_ENERGY.
 * qbf_set_moment_correction(e, mask) applies it inside qbf_inhom_step_periodic after the
 * collision (0 = off, the reference's behaviour). */
enum { QBF_CONS_MASS = 1, QBF_CONS_MOMENTUM = 2, QBF_CONS_ENERGY = 4 };
int qbf_set_moment_correction(qbf_engine *e, int mask);
int qbf_moment_correct(qbf_engine *e, const double *dWold, double *dWnew, int ncell, double L, int mask);

/* Per-cell diagnostics of physical states [ncell][4][N][N] into dOut[ncell][10] (device):
 * mass and spin densities (4), momentum (2), energy, entropy sum_v [s(l+) + s(l-)] dv with the binary
 * entropy s and the eigenvalues l+- = w0 +- |w| of w0 I + w.sigma, and the min and max eigenvalue
 * (visualization/wignerMoments.m:15-37, entropy.m, visualizeHomBoltzmann.m:26-45,98-105). */
int qbf_cell_diagnostics(qbf_engine *e, const double *dW, int ncell, double L, double *dOut);

/* Paired path (no counterpart in the reference; an option of the FP32 convolution path).  For
 * spectra of real states every grid that K_q = FFT2(psi2 .* IFFT2(.)) acts on is real, so one complex
 * application carries two, and the I1 and I2 chains of two components share one transform; the
 * imaginary parts this drops come from the Nyquist lines of the state (zero padding places them
 * without their conjugate partners; at N = 128 the I2 chains stay unpaired).  mode 0: off; 1: on when the
 * measured max |Im| / max |Re| of the padded physical grids of the state is at most tol; 2: always.
 * The paired result deviates from the complex one by about 0.15-0.8 times the ratio.  Defaults:
 * FP64 engines 0, FP32 engines 1 with tol = 2e-5.  qbf_real_pairs_status: whether the last collision evaluation used it, and the ratio
 * measured (-1 when not measured). */
int qbf_set_real_pairs(qbf_engine *e, int mode, double tol);
int qbf_real_pairs_status(qbf_engine *e, int *used, double *ratio);

/* Fermi-Dirac equilibrium with each cell's conserved moments (visualization/asymptoticFD.m,
 * calculateFDparams.m, createFD.m).  dParams[ncell][12] = beta, mu1, mu2, vel_x, vel_y, n_x, n_y, n_z
 * (eigenvector of the larger eigenvalue of the spin density matrix), lambda1 <= lambda2 (its
 * eigenvalues), intrinsic energy per particle, status (0 fitted, 1 no Fermi-Dirac state has these
 * moments).  dWFD (optional, may be NULL): [ncell][4][N][N] Pauli components of the fitted states. */
int qbf_fermi_dirac(qbf_engine *e, const double *dW, int ncell, double L, double *dParams, double *dWFD);

/* distance and quantum relative entropy to Fermi-Dirac states (visualizeHomBoltzmann.m:120-147,
 * visualizeInhomBoltzmann.m:163-205): dOut[ncell][2] = sum_j ||W_j - WFD_j||_1 dv (matrix 1-norm,
 * maximum over v_x of the sum over v_y), sum_v S(W(v) || WFD(v)) dv over WFD_0(v) > 1e-6.
 * fd_stride 0: one state dWFD for all cells; 1: one per cell.  N <= 64. */
int qbf_fd_distance(qbf_engine *e, const double *dW, int ncell, double L, const double *dWFD, int fd_stride,
	double *dOut);

/* average state over the cells, summed in cell order (visualizeInhomBoltzmann.m:80-89) */
int qbf_cell_average(qbf_engine *e, const double *dW, int ncell, double *dOut);

/* Timing of the dominant kernel (the I3 K-operator kernel): qbf_profile(e, 1) clears and
 * starts recording HIP events around each of its launches on the engine stream;
 * qbf_profile_read returns the summed kernel time, the launch count and the number of
 * K_q applications (IFFT2 -> multiply -> FFT2 on a 2N x 2N grid) they performed. */
int qbf_profile(qbf_engine *e, int enable);
int qbf_profile_read(qbf_engine *e, double *ms_total, int *launches, double *kapps);

/* quadrature tables in the reference's host layout, for runtime N (src/quadrature.c) */
void qbf_legendre_gauss(unsigned M, double a, double b, double *x, double *w);
void qbf_fourier_I1(int N, unsigned J, double L, double R, double *psiR1, double *psiR2);
void qbf_fourier_I2(int N, unsigned J, double L, double R, double *psiR1, double *psiR2);
void qbf_fourier_I3(int N, unsigned J, double L, double R, unsigned M, double *psiR1, double *psiR2);
void qbf_fourier_I4(int N, unsigned J, double L, double R, double *psiR1, double *psiR2);

/* ---- CNN surrogate of the collision operator (no counterpart in the reference) ----
 * Six periodic 3x3 convolutions 4 -> 64 -> 64 -> 64 -> 64 -> 64 -> 4 on N x N grids (N = 64 or
 * 128), GELU between them, bf16 weights and activations, fp32 accumulation on the matrix cores.
 * Inputs / outputs: planar fp32 [B][4][N][N] device arrays.  Parameters (host, fp32):
 * per layer weights [9][cout_pad][cin_pad] then bias [cout_pad]; cin_pad = 32 for layer 0
 * (4 used), cout_pad = 32 for layer 5 (4 used), 64 otherwise; unused entries are zero. */
typedef struct qbf_cnn qbf_cnn;
const char *qbf_cnn_last_error(void);
int    qbf_cnn_create(int N, int maxBatch, qbf_cnn **out);
void   qbf_cnn_destroy(qbf_cnn *c);
int    qbf_cnn_set_stream(qbf_cnn *c, void *hip_stream);
/* dilation per layer (dil[6], each in [1, 16]; NULL keeps the current ones) and per-cell
 * normalisation (norm = 1: the network sees x / a_b and the channel iscale * a_b, a_b = max |x[b][0]|;
 * its output is scaled by oscale * a_b^2 and the training loss is taken on the normalised output) */
int    qbf_cnn_configure(qbf_cnn *c, const int *dil, int norm);
size_t qbf_cnn_param_count(void);
int    qbf_cnn_set_params(qbf_cnn *c, const float *host);
int    qbf_cnn_get_params(qbf_cnn *c, float *host, int grads);
/* out = oscale * net(iscale * in) */
int    qbf_cnn_forward(qbf_cnn *c, const float *dIn, int B, float iscale, float *dOut, float oscale);
/* forward + backward of loss = mean((out - target)^2); gradients stay on the device;
 * B <= min(maxBatch, 256) (the training buffers are sized for that) */
int    qbf_cnn_gradients(qbf_cnn *c, const float *dIn, const float *dTarget, int B, float iscale, float oscale,
	float *dOut, float *hostLoss);
int    qbf_cnn_adam(qbf_cnn *c, float lr, float b1, float b2, float eps);
/* Midpoint collision step with C(W) ~ oscale * net(iscale * W) on physical states
 * [ncell][4][N][N] (device doubles, the engine's N = the network's), W_half = W + dt/2 C(W), out = W + dt C(W_half), then the
 * FP64 moment correction of out - W (mask QBF_CONS_*, 0 = none).  Launches on e's stream. */
int    qbf_surrogate_collision_step(qbf_cnn *c, qbf_engine *e, const double *dW, int ncell, double dt,
	float iscale, float oscale, double L, int mask, double *dOut);
/* Over the cells hostIdx[0..nsample-1] of W (physical, device), with C = Cc + Cd computed exactly by
 * e and C_net the network's: sums[0] = sum|C_net - C|, sums[1] = sum|C|, sums[2] = sum|W|;
 * nsample <= 64.  Synchronous. */
int    qbf_surrogate_residual(qbf_cnn *c, qbf_engine *e, const double *dW, int ncell, const int *hostIdx,
	int nsample, float iscale, float oscale, double sums[3]);
/* Offset field of the surrogate operator: C(W) ~ oscale * net(iscale * W) + D in the calls above and
 * below (D: fp32 planar [ncell][4][N][N] device array on the same cells as their states, kept by
 * pointer until the next call; NULL removes it).  While D is set, the stepping and residual calls
 * reject any other cell count.  With D = C(W0) - oscale * net(iscale * W0) at an exact step's
 * state W0 (qbf_last_collision) the surrogate follows the exact operator there and the network
 * supplies only its change. */
int    qbf_surrogate_set_offset(qbf_cnn *c, const float *dD, int ncell);
/* Strang step as qbf_inhom_step_periodic with the surrogate collision step; ncell <= maxBatch. */
int    qbf_surrogate_inhom_step_periodic(qbf_cnn *c, qbf_engine *e, const double *dW, int ncell, double L,
	double h, double dt, float iscale, float oscale, int mask, double *dOut);
/* that step captured as a HIP graph (qbf_graph_capture_calls; launch with qbf_graph_launch).  The
 * graph keeps the network's weights and the offset field by address: later weight updates and an
 * offset field rewritten in place are seen by replays; another offset array needs a new capture. */
int    qbf_graph_capture_surrogate_inhom_step_periodic(qbf_cnn *c, qbf_engine *e, const double *dW, int ncell,
	double L, double h, double dt, float iscale, float oscale, int mask, double *dOut, qbf_graph **g);

/* device memory helpers for host code that does not include HIP headers */
void *qbf_malloc(size_t bytes);
void  qbf_free(void *p);
int   qbf_h2d(qbf_engine *e, void *dst, const void *src, size_t bytes);
int   qbf_d2h(qbf_engine *e, void *dst, const void *src, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif
