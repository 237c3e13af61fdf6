// qbf_engine.hip -- host side of the MI355X collision-plus-transport engine (C API in include/qbf.h).
//
// The engine owns the device copies of the quadrature tables (reordered into the
// kernels' native layouts), the intermediate padded grids of one collision
// evaluation, and launches the kernel sequence:
//   What_c = FFT2(pad W_c)                                   k_padfft
//   Ahat_g = FFT2(sum_i psi1_i .* IFFT2(Fhat .* FFT2(pad(Phi_i .* W_g))))   k_i1acc  (Cc, I4)
//   S_i    = sum_q FFT2(pad(psi1_q W_i)) .* FFT2(pad(psi2_q W_i))           k_i2acc  (I2)
//   acc_j  = sum_q sum_i eta_i F_iq .* K_q[What_i .* F_jq]                   k_padfft + k_i3acc (I3)
//   C_j    = Extract(IFFT2( linear combination ))                           k_combine
// The reassociation is exact: the I1 weights of the reference are duplicated over
// one angle (src/quadrature.c:192-195, 213-234) so the J^2 terms are summed into J
// groups with Phi_i = sum_j psiR2_ij; the final convolution and Extract of every
// integral are linear and are hoisted out of the term sums.
#include "qbf_big.hip"
#include "qbf_i3mfma.hip"
#include "../../include/qbf.h"

#include <math.h>
#include <stdio.h>
#include <stdarg.h>
#include <string.h>
#include <deque>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>
#include <type_traits>

using namespace qbf;

static thread_local std::string g_err;

static int set_err(int code, const char* fmt, ...)
{
	char buf[512];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof(buf), fmt, ap);
	va_end(ap);
	g_err = buf;
	return code;
}

#define HIPCHK(x)                                                                              \
	do {                                                                                       \
		hipError_t e_ = (x);                                                                   \
		if (e_ != hipSuccess) return set_err(-2, "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
	} while (0)

extern "C" const char* qbf_last_error(void) { return g_err.c_str(); }

// set while a step is captured into a HIP graph (qbf_graph_capture_inhom_step_periodic): no
// allocation, host copy, read-back or synchronisation may happen then
static thread_local bool tl_capturing = false;
// set inside qbf_simulate_periodic_noalloc: a buffer that would have to grow, a combination table
// not yet on the device or a profiling event that would have to be created is an error instead
static thread_local bool tl_noalloc = false;

// The two state buffers of qbf_simulate_periodic_noalloc.  Each device holds its own copy of this
// array (allocated when the code object is loaded there, kept until the process ends); host code
// uses only the address that hipGetSymbolAddress returns for the device (g_scratch_dev), never the
// array's name.  Calls on one device share it, so calls that overlap on one device race.
static constexpr size_t QBF_NOALLOC_SCRATCH_BYTES = (size_t)QBF_NOALLOC_STATE_BYTES * 2;
static __device__ __align__(256) unsigned char scratch[QBF_NOALLOC_SCRATCH_BYTES];
static constexpr int QBF_MAX_DEVICES = 64;
static void* g_scratch_dev[QBF_MAX_DEVICES];
static std::mutex g_scratch_mu;

// the current device and its scratch address (looked up once per device)
static int scratch_for_current_device(int* devOut, void** ptrOut)
{
	int dev = -1;
	hipError_t st = hipGetDevice(&dev);
	if (st != hipSuccess || dev < 0 || dev >= QBF_MAX_DEVICES)
		return set_err(-2, "hipGetDevice failed or device %d beyond %d", dev, QBF_MAX_DEVICES);
	std::lock_guard<std::mutex> lk(g_scratch_mu);
	if (!g_scratch_dev[dev]) {
		void* p = nullptr;
		st = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (st != hipSuccess || !p)
			return set_err(-2, "hipGetSymbolAddress(scratch) on device %d failed: %s", dev, hipGetErrorString(st));
		g_scratch_dev[dev] = p;
	}
	*devOut = dev;
	*ptrOut = g_scratch_dev[dev];
	return 0;
}

// threads per workgroup for a padded P x P grid (P = 2N), and for an unpadded N x N grid
// XG: exchange rows and columns through a per-workgroup global scratch slot instead of LDS
// (an FP64 128 x 128 grid is 256 KiB, more than the 160 KiB of LDS of a CU)
template <typename R, int N> struct Cfg;
template <> struct Cfg<double, 16> { static constexpr int T = 128; static constexpr bool XG = false; };
template <> struct Cfg<double, 32> { static constexpr int T = 256; static constexpr bool XG = false; };
template <> struct Cfg<double, 64> { static constexpr int T = 1024; static constexpr bool XG = true; };
template <> struct Cfg<float, 16> { static constexpr int T = 128; static constexpr bool XG = false; };
template <> struct Cfg<float, 32> { static constexpr int T = 256; static constexpr bool XG = false; };
#ifndef QBF_F32_N64_THREADS
#define QBF_F32_N64_THREADS 1024   // 512 also fits in registers (G = 4, E = 32)
#endif
template <> struct Cfg<float, 64> { static constexpr int T = QBF_F32_N64_THREADS; static constexpr bool XG = false; };
// N = 128: the padded grids stay in global memory (qbf_big.hip: one workgroup of 512 threads per
// grid, line sweeps; T must match big::TB, XG sizes the scratch per grid)
template <> struct Cfg<double, 128> { static constexpr int T = 512; static constexpr bool XG = true; };
template <> struct Cfg<float, 128> { static constexpr int T = 512; static constexpr bool XG = true; };
template <int N> struct CfgNN;
template <> struct CfgNN<16> { static constexpr int T = 64; static constexpr bool XG = false; };
template <> struct CfgNN<32> { static constexpr int T = 128; static constexpr bool XG = false; };
template <> struct CfgNN<64> { static constexpr int T = 256; static constexpr bool XG = false; };
// an FP64 128 x 128 state grid (256 KiB) exchanges through global memory as well
template <> struct CfgNN<128> { static constexpr int T = 1024; static constexpr bool XG = true; };

struct qbf_graph {
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	hipStream_t stream = nullptr;    // the engine's stream at capture time; replays launch there
	struct qbf_engine* e = nullptr;         // null once the engine is destroyed (the graph is then dead)
	hipEvent_t before = nullptr, after = nullptr;   // order replays against the engine's current stream
	// the engine's record of its last periodic step (qbf_last_collision) as the captured calls left
	// it; a replay sets it again (setsLast) when the calls changed it
	bool setsLast = false;
	int lastCells = 0;
	double lastDt = 0.0;
	void release()
	{
		if (exec) (void)hipGraphExecDestroy(exec);
		if (graph) (void)hipGraphDestroy(graph);
		if (before) (void)hipEventDestroy(before);
		if (after) (void)hipEventDestroy(after);
		exec = nullptr;
		graph = nullptr;
		before = after = nullptr;
	}
};

struct qbf_engine {
	virtual ~qbf_engine()
	{
		if (sflux) (void)hipFree(sflux);
		for (auto& ev : evs) {
			hipEventDestroy(ev.first);
			hipEventDestroy(ev.second);
		}
	}
	int gridN = 0, precision = 0, maxCells = 0;
	int momentMask = 0;               // conservation correction after each collision step (0 = off)
	// paired I3 path: 0 off, 1 when the measured imaginary/real ratio of the padded physical grids
	// is at most pairTol, 2 always; pairRatio / pairUsed: the last collision evaluation's measurement
	// and choice
	int pairMode = 0;
	double pairTol = 0.0, pairRatio = -1.0;
	int pairUsed = 0;
	// optional timing of the I3 kernel (k_i3acc) launches with HIP events on the engine stream
	bool prof = false;
	std::vector<std::pair<hipEvent_t, hipEvent_t>> evs;
	size_t nev = 0;
	double profK = 0;                 // K_q applications issued while profiling
	int prof_begin(hipEvent_t* b, hipEvent_t* e)
	{
		if (nev == evs.size()) {
			if (tl_noalloc) return -1;

			hipEvent_t x, y;
			if (hipEventCreate(&x) != hipSuccess || hipEventCreate(&y) != hipSuccess) return -1;
			evs.push_back({x, y});
		}
		*b = evs[nev].first;
		*e = evs[nev].second;
		nev++;
		return 0;
	}
	hipStream_t stream = nullptr;
	bool ownStream = false;
	// the device the engine was created on and that device's scratch array (g_scratch_dev)
	int device = -1;
	void* scratchDev = nullptr;
	// HIP graphs captured from this engine hold raw pointers to its buffers and cached tables: while
	// any is alive (liveGraphs > 0) no buffer may grow (be reallocated) and no cached table may be
	// dropped; the engine's destruction invalidates the graphs still alive (graphs)
	int liveGraphs = 0;
	std::vector<qbf_graph*> graphs;
	std::vector<hipStream_t> retired;   // own streams replaced while a graph captured on them was alive
	virtual int set_i1(int slot, unsigned num, const double* p1, const double* p2) = 0;
	virtual int set_i2(unsigned num, const double* p1, const double* p2) = 0;
	virtual int set_i3(unsigned num, const double* p1, const double* p2) = 0;
	virtual int collision(const double2* W, int ncell, int mode, double2* out, const double2* base,
		const double2* Wcb, const double* Bext, double cf) = 0;
	virtual int collision_step(const double2* W, int ncell, double dt, const double* Bext, double2* Wnext) = 0;
	virtual int I1(int slot, const double2* f, const double2* g, const double2* h, double2* out) = 0;
	virtual int I2(const double2* f, const double2* g, const double2* h, double2* out) = 0;
	virtual int I3(const double2* f, const double2* g, const double2* h, double2* out) = 0;
	virtual int phys2fourier(const double* in, int ngrid, double2* out) = 0;
	virtual int fourier2phys(const double2* in, int ngrid, double* out) = 0;
	virtual int transport(const double* W, int ncell, double L, double h, double dt, double* out) = 0;
	virtual int inhom_step(const double* W, int ncell, double L, double h, double dt, const double* Bext, double* out) = 0;
	// the last periodic Strang step: its cell count and dt (0 cells: none since the last other step)
	int lastCells = 0;
	double lastDt = 0.0;
	unsigned lastSerial = 0;    // counts the updates of lastCells / lastDt (graph capture compares it)
	virtual int last_collision(int ncell, double* Win, double* C) = 0;
	virtual int inhom_step_dirichlet(const double* W, int ncell, double L, double h, double dt, const double* Bext,
		double* out) = 0;
	int transport_dirichlet(const double* W, int ncell, double L, double h, double dt, double* out)
	{
		if (ncell < 2) return set_err(-1, "Dirichlet transport needs at least 2 cells");
		const int chunk = ncell < 32 ? ncell : 32;
		const long long nthr = 4LL * gridN * gridN * ((ncell + chunk - 1) / chunk);
		hipLaunchKernelGGL(k_transport_dirichlet, dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, stream, W, out,
			gridN, ncell, chunk, h, dt, L);
		return hipGetLastError() == hipSuccess ? 0 : set_err(-2, "k_transport_dirichlet launch failed");
	}
	// Maxwell boundaries: trace fluxes of the current state into sflux, then the limiter with the
	// reflected ghost cells (WL, WR: the boundary states on the standard grid, not normalised)
	double* sflux = nullptr;
	virtual int inhom_step_maxwell(const double* W, int ncell, double L, double h, double dt, double lambda,
		const double* WL, const double* WR, const double* Bext, double* out) = 0;
	int transport_maxwell(const double* W, int ncell, double L, double h, double dt, double lambda,
		const double* WL, const double* WR, double* out)
	{
		if (ncell < 1) return set_err(-1, "Maxwell transport needs at least 1 cell");
		if (!sflux && hipMalloc(&sflux, 4 * sizeof(double)) != hipSuccess) {
			sflux = nullptr;
			return set_err(-3, "hipMalloc of the trace fluxes failed");
		}
		hipLaunchKernelGGL(k_maxwell_flux, dim3(1), dim3(256), 0, stream, W, gridN, ncell, L, WL, WR, sflux);
		if (hipGetLastError() != hipSuccess) return set_err(-2, "k_maxwell_flux launch failed");
		const int chunk = ncell < 32 ? ncell : 32;
		const long long nthr = 4LL * gridN * gridN * ((ncell + chunk - 1) / chunk);
		hipLaunchKernelGGL(k_transport_maxwell, dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, stream, W, out,
			gridN, ncell, chunk, h, dt, L, lambda, WL, WR, (const double*)sflux);
		return hipGetLastError() == hipSuccess ? 0 : set_err(-2, "k_transport_maxwell launch failed");
	}
	int moment_correct(const double* Wold, double* Wnew, int ncell, double L, int mask)
	{
		hipLaunchKernelGGL(k_moment_correct, dim3(ncell), dim3(256), 0, stream, Wold, Wnew, gridN, L, mask);
		hipError_t e = hipGetLastError();
		return e == hipSuccess ? 0 : -2;
	}
};

namespace {

struct DevBuf {
	void* p = nullptr;
	size_t bytes = 0;
	const int* lock = nullptr;     // the owner's live-graph count: no reallocation while it is non-zero
	int ensure(size_t n)
	{
		if (n <= bytes) return 0;
		if (tl_capturing) return set_err(-1, "buffer growth during graph capture");
		if (tl_noalloc)
			return set_err(-1, "qbf_simulate_periodic_noalloc: an engine buffer would have to grow to %zu bytes; "
				"run qbf_inhom_step_periodic once on the same cell count first", n);
		if (lock && *lock)
			return set_err(-1, "buffer growth (%zu bytes) while a captured graph uses the engine's buffers: "
				"destroy the graph first, or capture after a call of the largest size", n);
		if (p) hipFree(p);
		p = nullptr;
		bytes = 0;
		if (hipMalloc(&p, n) != hipSuccess) return set_err(-3, "hipMalloc(%zu) failed", n);
		bytes = n;
		return 0;
	}
	~DevBuf()
	{
		if (p) hipFree(p);
	}
	template <typename X> X* as() const { return (X*)p; }
};

double2* g_tw256 = nullptr;

int ensure_twiddles()
{
	if (g_tw256) return 0;
	double2 h[256];
	for (int j = 0; j < 256; j++) h[j] = make_double2(cos(2.0 * M_PI * j / 256.0), sin(2.0 * M_PI * j / 256.0));
	HIPCHK(hipMalloc(&g_tw256, sizeof(h)));
	HIPCHK(hipMemcpy(g_tw256, h, sizeof(h), hipMemcpyHostToDevice));
	return 0;
}

template <typename R, int N>
struct EngineT final : public qbf_engine {
	static constexpr int P = 2 * N, NN = N * N, PP = P * P;
	static constexpr int T = Cfg<R, N>::T;
	static constexpr int TN = CfgNN<N>::T;
	static constexpr bool XG = Cfg<R, N>::XG;
	static constexpr bool BIG = N == 128;        // P = 256 kernels of qbf_big.hip, natural grid order
	static_assert(!BIG || T == big::TB<R>, "launch size of the N = 128 kernels");
	using GE = Geo<P, T>;
	using C = cx<R>;
	static constexpr size_t GB = sizeof(C) * PP;    // bytes of one padded grid

	struct I1Tab { int n = 0; DevBuf phi, psi1n; };
	I1Tab i1[2];
	int n2 = 0; DevBuf i2p1, i2p2, i2pc;
	int n3 = 0; DevBuf i3p1, i3p2n;
	// BF16 hi/lo split path (precision QBF_BF16X3, N = 64): the I3 K-operator runs k_i3acc_mfma,
	// whose psi2 table has its own order (i3p2m)
	bool mfmaI3 = false;
	DevBuf i3p2m;

	DevBuf What, Ahat, Bhat, Sbuf, acc3, Fscr, xscr;
	// Combination-term tables: each distinct table is uploaded once and its device copy reused, so
	// that evaluations after the first of a configuration issue no host copy and no synchronisation
	// (which makes the step capturable in a HIP graph).  At most TERM_CACHE tables are kept.
	struct TermTab {
		std::vector<CTerm> host;
		DevBuf dev;
	};
	static constexpr size_t TERM_CACHE = 64;
	std::deque<TermTab> termCache;
	DevBuf Wmid, WF, WF2, Wt1, Wt2, src3;
	DevBuf Wpk, WhatP, reim, Ipk;     // paired path: packed spectra, their padded transforms, realness,
	                                  // packed I1-type results
	// device memory for the F grids of one I3 launch (it sets the quadrature terms per launch; the
	// accumulators cross HBM once per launch).  Measured at N = 64 on 512 cells: FP32 8 GiB (64 terms
	// per launch) 329 ms per step against 336 ms with 2 GiB; FP64 16 GiB no faster than 2 GiB.
	size_t fBudget = std::is_same<R, float>::value ? ((size_t)8 << 30) : ((size_t)2 << 30);

	int init()
	{
		for (DevBuf* b : {&i1[0].phi, &i1[0].psi1n, &i1[1].phi, &i1[1].psi1n, &i2p1, &i2p2, &i2pc, &i3p1, &i3p2n,
			     &What, &Ahat, &Bhat, &Sbuf, &acc3, &Fscr, &xscr, &Wmid, &WF, &WF2, &Wt1, &Wt2, &src3, &Wpk,
			     &WhatP, &reim, &Ipk, &xscrN, &Cf})
			b->lock = &liveGraphs;
		return ensure_twiddles();
	}

	// --------------------------------------------------------------- tables
	static void to_Ci_native(const double* tab, R* out)
	{
		if constexpr (BIG) {
			for (int i = 0; i < PP; i++) out[i] = (R)tab[i];
			return;
		}
		for (int t = 0; t < T; t++) {
			const int line = t / GE::G, p = t % GE::G;
			for (int s = 0; s < GE::E; s++) {
				const int l = p + GE::G * s, k = line;
				out[(size_t)s * T + t] = (R)tab[k + (size_t)P * l];
			}
		}
	}

	int upload(DevBuf& b, const void* h, size_t bytes)
	{
		int r = b.ensure(bytes);
		if (r) return r;
		HIPCHK(hipMemcpy(b.p, h, bytes, hipMemcpyHostToDevice));
		return 0;
	}

	int set_i1(int slot, unsigned num, const double* p1, const double* p2) override
	{
		if (slot < 0 || slot > 1) return set_err(-1, "bad I1 table slot %d", slot);
		// group terms whose psiR1 grids are identical: sum_{q in g} psi1 .* (f (*) (psiR2_q g))
		// = psi1 .* (f (*) ((sum_q psiR2_q) g)) exactly in exact arithmetic.
		std::vector<int> rep;                    // representative term of each group
		std::vector<std::vector<int>> members;
		std::unordered_map<uint64_t, std::vector<int>> byHash;
		for (unsigned q = 0; q < num; q++) {
			const double* g = p1 + (size_t)q * PP;
			uint64_t hsh = 1469598103934665603ULL;
			const uint64_t* w = (const uint64_t*)g;
			for (int i = 0; i < PP; i++) hsh = (hsh ^ w[i]) * 1099511628211ULL;
			int found = -1;
			auto it = byHash.find(hsh);
			if (it != byHash.end())
				for (int gi : it->second)
					if (memcmp(p1 + (size_t)rep[gi] * PP, g, sizeof(double) * PP) == 0) {
						found = gi;
						break;
					}
			if (found < 0) {
				found = (int)rep.size();
				rep.push_back((int)q);
				members.push_back({});
				byHash[hsh].push_back(found);
			}
			members[found].push_back((int)q);
		}
		const int ng = (int)rep.size();
		std::vector<double> phi((size_t)ng * NN, 0.0);
		std::vector<R> psi1n((size_t)ng * PP);
		for (int gi = 0; gi < ng; gi++) {
			double* d = &phi[(size_t)gi * NN];
			for (int q : members[gi])
				for (int e = 0; e < NN; e++) d[e] += p2[(size_t)q * NN + e];
			to_Ci_native(p1 + (size_t)rep[gi] * PP, &psi1n[(size_t)gi * PP]);
		}
		int r = upload(i1[slot].phi, phi.data(), sizeof(double) * phi.size());
		if (!r) r = upload(i1[slot].psi1n, psi1n.data(), sizeof(R) * psi1n.size());
		if (r) return r;
		i1[slot].n = ng;
		return 0;
	}

	int set_i2(unsigned num, const double* p1, const double* p2) override
	{
		int r = upload(i2p1, p1, sizeof(double) * NN * (size_t)num);
		if (!r) r = upload(i2p2, p2, sizeof(double) * NN * (size_t)num);
		if (r) return r;
		// psi1 + i psi2 for the paired I2 path (one transform yields both real factors)
		std::vector<double> pc(2 * NN * (size_t)num);
		for (size_t i = 0; i < NN * (size_t)num; i++) {
			pc[2 * i] = p1[i];
			pc[2 * i + 1] = p2[i];
		}
		r = upload(i2pc, pc.data(), sizeof(double) * pc.size());
		if (r) return r;
		n2 = (int)num;
		return 0;
	}

	int set_i3(unsigned num, const double* p1, const double* p2) override
	{
		std::vector<R> p2n((size_t)num * PP);
		for (unsigned q = 0; q < num; q++) to_Ci_native(p2 + (size_t)q * PP, &p2n[(size_t)q * PP]);
		int r = upload(i3p1, p1, sizeof(double2) * NN * (size_t)num);
		if (!r) r = upload(i3p2n, p2n.data(), sizeof(R) * p2n.size());
		if (r) return r;
		if (mfmaI3) {
			// k_i3acc_mfma's order: psi2 of row 16 w + i16, column 4 kq + r + 16 t at
			// ((w * 8 + t) * 4 + r) * 64 + 16 kq + i16
			static_assert(P == mf::P || N != 64, "the matrix-core kernel is built for P = 128");
			std::vector<float> p2m((size_t)num * PP);
			for (unsigned q = 0; q < num; q++)
				for (int w = 0; w < 8; w++)
					for (int t = 0; t < 8; t++)
						for (int rr = 0; rr < 4; rr++)
							for (int ln = 0; ln < 64; ln++) {
								const int l = 16 * w + (ln & 15), k = 4 * (ln >> 4) + rr + 16 * t;
								p2m[(size_t)q * PP + ((w * 8 + t) * 4 + rr) * 64 + ln] =
									(float)p2[(size_t)q * PP + k + (size_t)P * l];
							}
			r = upload(i3p2m, p2m.data(), sizeof(float) * p2m.size());
			if (r) return r;
		}
		n3 = (int)num;
		return 0;
	}

	// --------------------------------------------------------------- launches
	// exchange scratch for nblocks workgroups (only for configurations that exchange through global memory)
	int xs_for(int nblocks, void** xs, int gridsPerBlock = 1)
	{
		*xs = nullptr;
		if (!XG) return 0;
		const size_t per = BIG ? (size_t)PP * gridsPerBlock : (size_t)Xpose<R, GE>::SIZE;
		int r = xscr.ensure((size_t)nblocks * per * sizeof(C));
		if (r) return r;
		*xs = xscr.p;
		return 0;
	}

	int launch_padfft(const double2* src, int nsrc, const void* pre, int preKind, int q0, int nq, int nslot,
		const int* srcOfSlot, void* out, int ncell)
	{
		PadFFTArgs a{};
		a.src = src; a.nsrc = nsrc; a.pre = pre; a.preKind = preKind; a.q0 = q0; a.nq = nq; a.nslot = nslot;
		for (int i = 0; i < nslot; i++) a.srcOfSlot[i] = srcOfSlot[i];
		a.out = out; a.tw256 = g_tw256;
		a.ngrid = ncell * nq * nslot;
		const int nb = BIG ? a.ngrid : (a.ngrid + 1) / 2;
		int r = xs_for(nb, &a.xs);
		if (r) return r;
		if constexpr (BIG) {
			hipLaunchKernelGGL((k_padfft_big<R>), dim3(nb), dim3(T), 0, stream, a);
		} else if constexpr (!XG && N == 64 && std::is_same<R, float>::value) {
			// one workgroup per CU, each looping over pairs of grids (measured: FP32 step 354 -> 335 ms at
			// Nx = 512; 256 and 512 workgroups alike)
			constexpr int WG = 256;
			hipLaunchKernelGGL((k_padfft_loop<R, N, T>), dim3(nb < WG ? nb : WG), dim3(T), 0, stream, a);
		} else {
			hipLaunchKernelGGL((k_padfft<R, N, T, XG>), dim3(nb), dim3(T), 0, stream, a);
		}
		HIPCHK(hipGetLastError());
		return 0;
	}

	// Term ranges per (cell, instance) so that a launch has about 1024 workgroups when there are few
	// cells, each range keeping at least minTerms terms; the partial sums land in consecutive slots.
	static int term_chunks(int ncell, int ninst, int nterm, int minTerms)
	{
		int c = (1024 + ncell * ninst - 1) / (ncell * ninst);
		const int cmax = nterm / minTerms;
		if (c > cmax) c = cmax;
		return c < 1 ? 1 : c;
	}

	int launch_i1(int slot, const double2* src, int nsrc, const void* Fhat, int nFh, double fa, double fb,
		int ninst, const int* instSrc, const int* instFh, DevBuf& out, int ncell, int* nchunk)
	{
		if (i1[slot].n <= 0) return set_err(-1, "I1-type table slot %d not set", slot);
		const int nch = term_chunks(ncell, ninst, i1[slot].n, 4);
		int r = out.ensure((size_t)ncell * ninst * nch * GB);
		if (r) return r;
		I1Args a{};
		a.src = src; a.nsrc = nsrc; a.Fhat = Fhat; a.nFh = nFh; a.fa = fa; a.fb = fb;
		a.phi = i1[slot].phi.template as<double>(); a.psi1n = i1[slot].psi1n.p; a.nterm = i1[slot].n;
		a.ninst = ninst;
		for (int i = 0; i < ninst; i++) { a.instSrc[i] = instSrc[i]; a.instFh[i] = instFh[i]; }
		a.nchunk = nch;
		a.out = out.p; a.tw256 = g_tw256;
		r = xs_for(ncell * ninst * nch, &a.xs);
		if (r) return r;
		if constexpr (BIG)
			hipLaunchKernelGGL((k_i1acc_big<R>), dim3(ncell * ninst * nch), dim3(T), 0, stream, a);
		else
			hipLaunchKernelGGL((k_i1acc<R, N, T, XG>), dim3(ncell * ninst * nch), dim3(T), 0, stream, a);
		HIPCHK(hipGetLastError());
		*nchunk = nch;
		return 0;
	}

	// pair: f = g for every instance and the grids are real (see k_i2acc, PAIR)
	int launch_i2(const double2* src, int nsrc, int ninst, const int* instF, const int* instG, DevBuf& out, int ncell,
		int* nchunk, bool pair = false)
	{
		if (n2 <= 0) return set_err(-1, "I2 table not set");
		if (BIG && pair) return set_err(-1, "no paired I2 kernel at N = 128");
		if (pair)
			for (int i = 0; i < ninst; i++)
				if (instF[i] != instG[i]) return set_err(-1, "paired I2 needs f = g");
		const int nch = term_chunks(ncell, ninst, n2, 2);
		int r = out.ensure((size_t)ncell * ninst * nch * GB);
		if (r) return r;
		I2Args a{};
		a.src = src; a.nsrc = nsrc; a.psi1 = i2p1.template as<double>(); a.psi2 = i2p2.template as<double>();
		a.psic = i2pc.template as<double2>();
		a.nterm = n2; a.ninst = ninst;
		for (int i = 0; i < ninst; i++) { a.instF[i] = instF[i]; a.instG[i] = instG[i]; }
		a.nchunk = nch;
		a.out = out.p; a.tw256 = g_tw256;
		r = xs_for(ncell * ninst * nch, &a.xs);
		if (r) return r;
		if constexpr (BIG)
			hipLaunchKernelGGL((k_i2acc_big<R>), dim3(ncell * ninst * nch), dim3(T), 0, stream, a);
		else if (pair)
			hipLaunchKernelGGL((k_i2acc<R, N, T, XG, true>), dim3(ncell * ninst * nch), dim3(T), 0, stream, a);
		else
			hipLaunchKernelGGL((k_i2acc<R, N, T, XG, false>), dim3(ncell * ninst * nch), dim3(T), 0, stream, a);
		HIPCHK(hipGetLastError());
		*nchunk = nch;
		return 0;
	}

	// I3 over all quadrature terms; returns the number of accumulator chunks per (cell, out)
	int run_i3(const double2* src, int nsrc, int nF, const int* srcOfF, const void* Shat, int nS, int nout,
		const int* ninst, const int (*iF)[4], const int (*iS)[4], const int (*iH)[4], const double (*iW)[4],
		int ncell, int* nchunkOut, bool pair = false)
	{
		if (n3 <= 0) return set_err(-1, "I3 table not set");
		const int nq = n3;
		// term ranges: about 1024 workgroups when there are few cells, a power of two so that the
		// ranges of a power-of-two term count are equal and the workgroups fill whole waves
		int nchunk = 1;
		while (nchunk * ncell * nout < 1024) nchunk *= 2;
		if (nchunk > nq) nchunk = nq;
		if (nchunk < 1) nchunk = 1;
		const size_t perQ = (size_t)ncell * nF * GB * (XG ? 2 : 1);
		int qr = (int)(fBudget / perQ);
		if (qr < 1) qr = 1;
		if (qr > nq) qr = nq;
		if (nchunk > qr) nchunk = qr;
		int r = Fscr.ensure((size_t)ncell * nF * GB * qr);
		if (!r) r = acc3.ensure((size_t)ncell * nout * nchunk * GB);
		if (r) return r;
		for (int q0 = 0; q0 < nq; q0 += qr) {
			const int nqr = (nq - q0) < qr ? (nq - q0) : qr;
			r = launch_padfft(src, nsrc, i3p1.p, 2, q0, nqr, nF, srcOfF, Fscr.p, ncell);
			if (r) return r;
			I3Args a{};
			a.Shat = Shat; a.nS = nS; a.F = Fscr.p; a.nF = nF; a.nq = nqr; a.q0 = q0;
			a.psi2n = i3p2n.p; a.nout = nout;
			for (int o = 0; o < nout; o++) {
				a.ninst[o] = ninst[o];
				for (int i = 0; i < ninst[o]; i++) {
					a.instF[o][i] = iF[o][i]; a.instS[o][i] = iS[o][i]; a.instH[o][i] = iH[o][i];
				}
				// paired instances carry two weights each
				for (int i = 0; i < (pair ? 2 * ninst[o] : ninst[o]); i++) a.instW[o][i] = iW[o][i];
			}
			a.nchunk = nchunk; a.accumulate = q0 > 0; a.acc = acc3.p; a.tw256 = g_tw256;
			r = xs_for(ncell * nout * nchunk, &a.xs, BIG ? big::I3_NI : 1);
			if (r) return r;
			hipEvent_t eb = nullptr, ee = nullptr;
			if (prof && !tl_capturing && prof_begin(&eb, &ee) == 0) HIPCHK(hipEventRecord(eb, stream));
			if constexpr (BIG) {
				if (pair)
					hipLaunchKernelGGL((k_i3acc_big<R, true>), dim3(ncell * nout * nchunk), dim3(T), 0, stream, a);
				else
					hipLaunchKernelGGL((k_i3acc_big<R, false>), dim3(ncell * nout * nchunk), dim3(T), 0, stream, a);
			} else if (mfmaI3) {
				if constexpr (std::is_same<R, float>::value && N == 64 && T == 1024) {
					a.psi2n = i3p2m.p;
					if (pair)
						hipLaunchKernelGGL((k_i3acc_mfma<true>), dim3(ncell * nout * nchunk), dim3(mf::T), 0, stream, a);
					else
						hipLaunchKernelGGL((k_i3acc_mfma<false>), dim3(ncell * nout * nchunk), dim3(mf::T), 0, stream, a);
				} else {
					return set_err(-1, "the BF16 split path exists for N = 64 only");
				}
			} else if (pair)
				hipLaunchKernelGGL((k_i3acc<R, N, T, XG, true>), dim3(ncell * nout * nchunk), dim3(T), 0, stream, a);
			else
				hipLaunchKernelGGL((k_i3acc<R, N, T, XG, false>), dim3(ncell * nout * nchunk), dim3(T), 0, stream, a);
			HIPCHK(hipGetLastError());
			if (ee) {
				HIPCHK(hipEventRecord(ee, stream));
				double nk = 0;
				for (int o = 0; o < nout; o++) nk += ninst[o];
				profK += nk * ncell * nqr;
			}
		}
		*nchunkOut = nchunk;
		return 0;
	}

	// max |Im| / max |Re| over the padded physical grids of the state (What): the size of what the
	// paired I3 path drops, relative to what it keeps
	int measure_realness(int ncell, double* ratio)
	{
		constexpr int NB = 256;
		int r = reim.ensure(sizeof(double) * 2 * NB);
		if (r) return r;
		hipLaunchKernelGGL((k_re_im_max<R>), dim3(NB), dim3(256), 0, stream, What.template as<C>(),
			(long long)ncell * 4 * PP, reim.template as<double>());
		HIPCHK(hipGetLastError());
		double h[2 * NB];
		HIPCHK(hipMemcpyAsync(h, reim.p, sizeof(h), hipMemcpyDeviceToHost, stream));
		HIPCHK(hipStreamSynchronize(stream));
		double mr = 0, mi = 0;
		for (int b = 0; b < NB; b++) {
			mr = fmax(mr, h[2 * b]);
			mi = fmax(mi, h[2 * b + 1]);
		}
		*ratio = mr > 0 ? mi / mr : (mi > 0 ? INFINITY : 0.0);
		return 0;
	}

	// Packed sources of the paired path, [ncell][3]: W0 + i W1, W2 + i W3, W1 (see k_pack_pairs)
	int pack_sources(const double2* W, int ncell)
	{
		int r = Wpk.ensure(sizeof(double2) * 3 * NN * (size_t)ncell);
		if (r) return r;
		const long long n = 3LL * NN * ncell;
		hipLaunchKernelGGL(k_pack_pairs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, W,
			Wpk.template as<double2>(), NN, n);
		HIPCHK(hipGetLastError());
		return 0;
	}

	// Real parts and imaginary parts of packed I1-type results [ncell][npk][nch] into the slots
	// dst[k][0], dst[k][1] of out [ncell][nout][nch] (imaginary part zero); -1: not used.
	int unpack_real(const DevBuf& packed, int npk, DevBuf& out, int nout, int nch, const int (*dst)[2], int ncell)
	{
		int r = out.ensure((size_t)ncell * nout * nch * GB);
		if (r) return r;
		UnpackArgs a{};
		a.in = packed.p; a.out = out.p; a.npk = npk; a.nout = nout; a.nch = nch;
		for (int k = 0; k < npk; k++) { a.dst[k][0] = dst[k][0]; a.dst[k][1] = dst[k][1]; }
		const long long n = (long long)ncell * npk * nch * PP;
		hipLaunchKernelGGL((k_unpack_real<R>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, a, n, (long long)PP);
		HIPCHK(hipGetLastError());
		return 0;
	}

	// I3 part of CdInt on packed component pairs (see k_i3acc, PAIR); Wpk from pack_sources
	int run_i3_pairs(int ncell, int* nchunk)
	{
		int r = WhatP.ensure((size_t)ncell * 2 * GB);
		if (r) return r;
		const int two[2] = {0, 1};
		r = launch_padfft(Wpk.template as<double2>(), 3, nullptr, 0, 0, 1, 2, two, WhatP.p, ncell);
		if (r) return r;
		static const double eta[4] = {1, -1, -1, -1};
		int ni[4] = {2, 2, 2, 2};
		int iF[4][4] = {{0}}, iS[4][4] = {{0}}, iH[4][4] = {{0}};
		double iW[4][4] = {{0}};
		for (int j = 0; j < 4; j++)
			for (int pr = 0; pr < 2; pr++) {
				iF[j][pr] = pr;
				iS[j][pr] = pr;
				iH[j][pr] = 2 * (j >> 1) + (j & 1);
				iW[j][2 * pr] = eta[2 * pr];
				iW[j][2 * pr + 1] = eta[2 * pr + 1];
			}
		return run_i3(Wpk.template as<double2>(), 3, 2, two, WhatP.p, 2, 4, ni, iF, iS, iH, iW, ncell, nchunk, true);
	}

	int launch_combine(std::vector<CTerm> (&tl)[4], int nout, const int* outComp, int ncell, double2* out,
		const double2* base, const double2* Wcb, const double* Bext, double cf)
	{
		int maxT = 1;
		for (int o = 0; o < nout; o++) maxT = (int)tl[o].size() > maxT ? (int)tl[o].size() : maxT;
		std::vector<CTerm> flat((size_t)nout * maxT);
		for (int o = 0; o < nout; o++)
			for (size_t i = 0; i < tl[o].size(); i++) flat[(size_t)o * maxT + i] = tl[o][i];
		const CTerm* dterms = nullptr;
		int r = term_table(flat, &dterms);
		if (r) return r;
		CombineArgs a{};
		a.terms = dterms; a.maxTerms = maxT; a.nout = nout;
		for (int o = 0; o < nout; o++) { a.nterm[o] = (int)tl[o].size(); a.outComp[o] = outComp[o]; }
		a.outScale = 1.0; a.out = out; a.base = base; a.Wcb = Bext ? Wcb : nullptr; a.Bext = Bext; a.cf = cf;
		a.tw256 = g_tw256;
		r = xs_for(ncell * nout, &a.xs);
		if (r) return r;
		if constexpr (BIG)
			hipLaunchKernelGGL((k_combine_big<R>), dim3(ncell * nout), dim3(T), 0, stream, a);
		else
			hipLaunchKernelGGL((k_combine<R, N, T, XG>), dim3(ncell * nout), dim3(T), 0, stream, a);
		HIPCHK(hipGetLastError());
		return 0;
	}

	static bool same_terms(const std::vector<CTerm>& a, const std::vector<CTerm>& b)
	{
		if (a.size() != b.size()) return false;
		for (size_t i = 0; i < a.size(); i++) {
			const CTerm &x = a[i], &y = b[i];
			if (x.X != y.X || x.xcell != y.xcell || x.Y != y.Y || x.ycell != y.ycell || x.ysum != y.ysum ||
				x.nsum != y.nsum || x.coef != y.coef || x.ax != y.ax || x.bx != y.bx)
				return false;
		}
		return true;
	}

	// device copy of a combination-term table (uploaded on first use; the table's layout [nout][maxT]
	// is implied by its size and the caller's maxT, which the kernel receives separately)
	int term_table(const std::vector<CTerm>& flat, const CTerm** dev)
	{
		for (auto& t : termCache)
			if (same_terms(t.host, flat)) {
				*dev = t.dev.template as<CTerm>();
				return 0;
			}
		if (tl_capturing) return set_err(-1, "combination table not prepared: run the step once before capturing it");
		if (tl_noalloc)
			return set_err(-1, "qbf_simulate_periodic_noalloc: combination table not on the device; run "
				"qbf_inhom_step_periodic once on the same cell count first");
		if (termCache.size() >= TERM_CACHE && liveGraphs)
			return set_err(-1, "combination-table cache full while a captured graph uses its tables: destroy the graph first");
		if (termCache.size() >= TERM_CACHE) {
			HIPCHK(hipStreamSynchronize(stream));     // no kernel reads the table that is dropped
			termCache.pop_front();
		}
		termCache.emplace_back();
		TermTab& t = termCache.back();
		t.host = flat;
		int r = t.dev.ensure(sizeof(CTerm) * flat.size());
		if (r) {
			termCache.pop_back();
			return r;
		}
		HIPCHK(hipMemcpyAsync(t.dev.p, t.host.data(), sizeof(CTerm) * flat.size(), hipMemcpyHostToDevice, stream));
		*dev = t.dev.template as<CTerm>();
		return 0;
	}

	// buf[o][c] (+)= ..., o < nouter, c < nch grids: buf[o][0] = sum_c buf[o][c]
	int sum_chunks(void* buf, long long nouter, int nch)
	{
		if (nch <= 1) return 0;
		const long long n = nouter * PP;
		hipLaunchKernelGGL((k_sum_chunks<R>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, (C*)buf, nch,
			nouter, (long long)PP);
		HIPCHK(hipGetLastError());
		return 0;
	}

	static CTerm term(const void* X, long long xcell, const void* Y, long long ycell, double coef,
		double ax = 1, double bx = 0, int nsum = 1, long long ysum = 0)
	{
		CTerm t{};
		t.X = X; t.xcell = xcell; t.Y = Y; t.ycell = ycell; t.coef = coef; t.ax = ax; t.bx = bx;
		t.nsum = nsum; t.ysum = ysum;
		return t;
	}

	// --------------------------------------------------------------- collision operator
	int collision(const double2* W, int ncell, int mode, double2* out, const double2* base,
		const double2* Wcb, const double* Bext, double cf) override
	{
		if (ncell < 1 || ncell > maxCells) return set_err(-1, "ncell %d outside [1, %d]", ncell, maxCells);
		const bool cc = mode & QBF_CC, cd = mode & QBF_CD;
		const long long c4 = 4LL * PP;
		int r = What.ensure((size_t)ncell * 4 * GB);
		if (r) return r;
		int nchA = 1, nchB = 1, nchS = 1;
		const int all4[4] = {0, 1, 2, 3}, zero4[4] = {0, 0, 0, 0};
		r = launch_padfft(W, 4, nullptr, 0, 0, 1, 4, all4, What.p, ncell);
		if (r) return r;
		// paired path: every I1, I2 and I3 chain of a real state is real (see k_i3acc, k_i2acc); the I1
		// chain is complex-linear in its source, so W_a + i W_b yields out_a + i out_b
		pairUsed = 0;
		pairRatio = -1.0;
		if (pairMode == 1 && tl_capturing)
			return set_err(-1, "graph capture needs a fixed paired-path choice (qbf_set_real_pairs mode 0 or 2)");
		if (pairMode == 1) {
			r = measure_realness(ncell, &pairRatio);
			if (r) return r;
		}
		const bool pair = pairMode == 2 || (pairMode == 1 && pairRatio <= pairTol);
		if (pair) {
			pairUsed = 1;
			r = pack_sources(W, ncell);
			if (r) return r;
		}
		if (cc) {
			// -2 (2 w_tr - 1) ... : Fhat = FFT2(pad(-4 W0 + 2 delta)) = -4 What_0 + 2 (src/collision.c:220-227)
			if (pair) {
				// packed sources W2 + i W3 and W1 -> Ahat slots (1, 2) and (0, -)
				const int srcs[2] = {1, 2}, dst[2][2] = {{1, 2}, {0, -1}};
				r = launch_i1(QBF_TAB_I1, Wpk.template as<double2>(), 3, What.p, 4, -4.0, 2.0, 2, srcs, zero4, Ipk,
					ncell, &nchA);
				if (!r) r = unpack_real(Ipk, 2, Ahat, 3, nchA, dst, ncell);
			} else {
				const int g3[3] = {1, 2, 3};
				r = launch_i1(QBF_TAB_I1, W, 4, What.p, 4, -4.0, 2.0, 3, g3, zero4, Ahat, ncell, &nchA);
			}
			if (r) return r;
		}
		int nchunk = 1;
		if (cd) {
			// I4 with Fhat = FFT2(pad(2 W0 - delta)) = 2 What_0 - 1 (src/collision.c:127-134)
			if (pair) {
				// packed sources W0 + i W1 and W2 + i W3 -> Bhat slots (0, 1) and (2, 3)
				const int srcs[2] = {0, 1}, dst[2][2] = {{0, 1}, {2, 3}};
				r = launch_i1(QBF_TAB_I4, Wpk.template as<double2>(), 3, What.p, 4, 2.0, -1.0, 2, srcs, zero4, Ipk,
					ncell, &nchB);
				if (!r) r = unpack_real(Ipk, 2, Bhat, 4, nchB, dst, ncell);
			} else {
				r = launch_i1(QBF_TAB_I4, W, 4, What.p, 4, 2.0, -1.0, 4, all4, zero4, Bhat, ncell, &nchB);
			}
			// (no paired I2 kernel at N = 128: its I2 runs on the complex sources)
			if (!r) r = launch_i2(W, 4, 4, all4, all4, Sbuf, ncell, &nchS, pair && !BIG);
			if (r) return r;
			if (pair) {
				r = run_i3_pairs(ncell, &nchunk);
			} else {
				static const double eta[4] = {1, -1, -1, -1};
				int ni[4] = {4, 4, 4, 4};
				int iF[4][4], iS[4][4], iH[4][4];
				double iW[4][4];
				for (int j = 0; j < 4; j++)
					for (int i = 0; i < 4; i++) { iF[j][i] = i; iS[j][i] = i; iH[j][i] = j; iW[j][i] = eta[i]; }
				r = run_i3(W, 4, 4, all4, What.p, 4, 4, ni, iF, iS, iH, iW, ncell, &nchunk);
			}
			if (r) return r;
		}
		// the partial sums over term ranges are added here over the whole grid, so that k_combine
		// reads one grid per term (added in the order k_combine used: the result is unchanged)
		if (cc) r = sum_chunks(Ahat.p, 3LL * ncell, nchA);
		if (!r && cd) r = sum_chunks(Bhat.p, 4LL * ncell, nchB);
		if (!r && cd) r = sum_chunks(Sbuf.p, 4LL * ncell, nchS);
		if (!r && cd) r = sum_chunks(acc3.p, 4LL * ncell, nchunk);
		if (r) return r;
		const double p2 = 1.0 / ((double)P * P), p4 = p2 * p2;
		const double eta2[4] = {2, -2, -2, -2};
		const C* Wh = What.template as<C>();
		auto Wslot = [&](int c) { return (const void*)(Wh + (size_t)c * PP); };
		std::vector<CTerm> tl[4];
		if (cc) {
			const C* A = Ahat.template as<C>();
			const long long a3 = 3LL * nchA * PP;
			auto As = [&](int s) { return (const void*)(A + (size_t)s * nchA * PP); };
			// Cc_1 = I1(f,W2,W3) - I1(f,W3,W2), cyclic (src/collision.c:231-253); Ahat slot s <-> g = s+1
			// (each slot: nchA partial sums over term ranges, added into the first)
			tl[1].push_back(term(Wslot(3), c4, As(1), a3, p4, 1, 0, 1, PP));
			tl[1].push_back(term(Wslot(2), c4, As(2), a3, -p4, 1, 0, 1, PP));
			tl[2].push_back(term(Wslot(1), c4, As(2), a3, p4, 1, 0, 1, PP));
			tl[2].push_back(term(Wslot(3), c4, As(0), a3, -p4, 1, 0, 1, PP));
			tl[3].push_back(term(Wslot(2), c4, As(0), a3, p4, 1, 0, 1, PP));
			tl[3].push_back(term(Wslot(1), c4, As(1), a3, -p4, 1, 0, 1, PP));
		}
		if (cd) {
			const C* B = Bhat.template as<C>();
			const C* S = Sbuf.template as<C>();
			const C* A3 = acc3.template as<C>();
			const long long b4 = 4LL * nchB * PP, s4 = 4LL * nchS * PP;
			auto Bs = [&](int s) { return (const void*)(B + (size_t)s * nchB * PP); };
			for (int j = 0; j < 4; j++) {
				// I4 terms (src/collision.c:137-167)
				tl[j].push_back(term(Wslot(0), c4, Bs(j), b4, eta2[j] * p4, 1, 0, 1, PP));
				tl[j].push_back(term(Wslot(j), c4, Bs(0), b4, 2 * p4, 1, 0, 1, PP));
				if (j == 0)
					for (int jj = 0; jj < 4; jj++)
						tl[j].push_back(term(Wslot(jj), c4, Bs(jj), b4, -2 * p4, 1, 0, 1, PP));
				// I3 terms: -eta2_j / P^4 * sum_chunks acc_j (src/collision.c:110-121)
				tl[j].push_back(term(nullptr, 0, A3 + (size_t)j * nchunk * PP, 4LL * nchunk * PP, -eta2[j] * p4,
					1, 0, 1, PP));
				// I2 terms: eta2_i / P^2 * (delta_j0 - What_j) .* S_i (src/collision.c:82-106)
				for (int i = 0; i < 4; i++)
					tl[j].push_back(term(Wslot(j), c4, S + (size_t)i * nchS * PP, s4, eta2[i] * p2, -1.0,
						j == 0 ? 1.0 : 0.0, 1, PP));
			}
		}
		return launch_combine(tl, 4, all4, ncell, out, base, Wcb, Bext, cf);
	}

	int collision_step(const double2* W, int ncell, double dt, const double* Bext, double2* Wnext) override
	{
		lastCells = 0;      // Wmid is rewritten (inhom_step sets it again afterwards)
		lastSerial++;
		int r = Wmid.ensure(sizeof(double2) * 4 * NN * (size_t)ncell);
		if (r) return r;
		double2* Wm = Wmid.template as<double2>();
		r = collision(W, ncell, QBF_C, Wm, W, W, Bext, 0.5 * dt);
		if (!r) r = collision(Wm, ncell, QBF_C, Wnext, W, Wm, Bext, dt);
		return r;
	}

	// --------------------------------------------------------------- single integrals (one grid)
	int stage3(const double2* f, const double2* g, const double2* h)
	{
		int r = src3.ensure(sizeof(double2) * 3 * NN);
		if (r) return r;
		double2* s = src3.template as<double2>();
		HIPCHK(hipMemcpyAsync(s, f, sizeof(double2) * NN, hipMemcpyDeviceToDevice, stream));
		HIPCHK(hipMemcpyAsync(s + NN, g, sizeof(double2) * NN, hipMemcpyDeviceToDevice, stream));
		HIPCHK(hipMemcpyAsync(s + 2 * NN, h, sizeof(double2) * NN, hipMemcpyDeviceToDevice, stream));
		return 0;
	}

	int I1(int slot, const double2* f, const double2* g, const double2* h, double2* out) override
	{
		int r = stage3(f, g, h);
		if (!r) r = What.ensure(2 * GB);
		if (r) return r;
		const double2* s = src3.template as<double2>();
		const int fh[2] = {0, 2}, one[1] = {1}, z[1] = {0};
		int nch = 1;
		r = launch_padfft(s, 3, nullptr, 0, 0, 1, 2, fh, What.p, 1);
		if (!r) r = launch_i1(slot, s, 3, What.p, 2, 1.0, 0.0, 1, one, z, Ahat, 1, &nch);
		if (r) return r;
		const double p4 = 1.0 / ((double)P * P * P * P);
		std::vector<CTerm> tl[4];
		tl[0].push_back(term(What.template as<C>() + PP, 0, Ahat.p, 0, p4, 1, 0, nch, PP));
		const int c0[1] = {0};
		return launch_combine(tl, 1, c0, 1, out, nullptr, nullptr, nullptr, 1.0);
	}

	int I2(const double2* f, const double2* g, const double2* h, double2* out) override
	{
		int r = stage3(f, g, h);
		if (!r) r = What.ensure(GB);
		if (r) return r;
		const double2* s = src3.template as<double2>();
		const int hh[1] = {2}, f0[1] = {0}, g1[1] = {1};
		int nch = 1;
		r = launch_padfft(s, 3, nullptr, 0, 0, 1, 1, hh, What.p, 1);
		if (!r) r = launch_i2(s, 3, 1, f0, g1, Sbuf, 1, &nch);
		if (r) return r;
		const double p2 = 1.0 / ((double)P * P);
		std::vector<CTerm> tl[4];
		tl[0].push_back(term(What.p, 0, Sbuf.p, 0, p2, 1, 0, nch, PP));
		const int c0[1] = {0};
		return launch_combine(tl, 1, c0, 1, out, nullptr, nullptr, nullptr, 1.0);
	}

	int I3(const double2* f, const double2* g, const double2* h, double2* out) override
	{
		int r = stage3(f, g, h);
		if (!r) r = What.ensure(GB);
		if (r) return r;
		const double2* s = src3.template as<double2>();
		const int gg[1] = {1}, fh[2] = {0, 2};
		r = launch_padfft(s, 3, nullptr, 0, 0, 1, 1, gg, What.p, 1);
		if (r) return r;
		int ni[4] = {1, 0, 0, 0};
		int iF[4][4] = {{0}}, iS[4][4] = {{0}}, iH[4][4] = {{1}};
		double iW[4][4] = {{1.0}};
		int nchunk = 1;
		r = run_i3(s, 3, 2, fh, What.p, 1, 1, ni, iF, iS, iH, iW, 1, &nchunk);
		if (r) return r;
		const double p4 = 1.0 / ((double)P * P * P * P);
		std::vector<CTerm> tl[4];
		tl[0].push_back(term(nullptr, 0, acc3.p, 0, p4, 1, 0, nchunk, PP));
		const int c0[1] = {0};
		return launch_combine(tl, 1, c0, 1, out, nullptr, nullptr, nullptr, 1.0);
	}

	// --------------------------------------------------------------- state transforms and transport
	// N x N state transforms; with XGN the exchange goes through xscrN, at most NB_XN grids a launch
	static constexpr bool XGN = CfgNN<N>::XG;
	static constexpr int NB_XN = 2048;
	DevBuf xscrN;
	int phys2fourier(const double* in, int ngrid, double2* out) override
	{
		for (int g0 = 0; g0 < ngrid; g0 += NB_XN) {
			const int ng = ngrid - g0 < NB_XN ? ngrid - g0 : NB_XN;
			double2* xs = nullptr;
			if (XGN) {
				int r = xscrN.ensure(sizeof(double2) * (size_t)NB_XN * Xpose<double, Geo<N, TN>>::SIZE);
				if (r) return r;
				xs = xscrN.template as<double2>();
			}
			hipLaunchKernelGGL((k_phys2fourier<N, TN, XGN>), dim3(ng), dim3(TN), 0, stream, in + (size_t)g0 * NN,
				out + (size_t)g0 * NN, g_tw256, ng, xs);
			HIPCHK(hipGetLastError());
		}
		return 0;
	}

	int fourier2phys(const double2* in, int ngrid, double* out) override
	{
		for (int g0 = 0; g0 < ngrid; g0 += NB_XN) {
			const int ng = ngrid - g0 < NB_XN ? ngrid - g0 : NB_XN;
			double2* xs = nullptr;
			if (XGN) {
				int r = xscrN.ensure(sizeof(double2) * (size_t)NB_XN * Xpose<double, Geo<N, TN>>::SIZE);
				if (r) return r;
				xs = xscrN.template as<double2>();
			}
			hipLaunchKernelGGL((k_fourier2phys<N, TN, XGN>), dim3(ng), dim3(TN), 0, stream, in + (size_t)g0 * NN,
				out + (size_t)g0 * NN, g_tw256, ng, xs);
			HIPCHK(hipGetLastError());
		}
		return 0;
	}

	int transport(const double* W, int ncell, double L, double h, double dt, double* out) override
	{
		if (ncell < 1) return set_err(-1, "ncell must be >= 1");
		const int chunk = ncell < 32 ? ncell : 32;
		const long long nthr = 4LL * NN * ((ncell + chunk - 1) / chunk);
		hipLaunchKernelGGL(k_transport_periodic, dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, stream,
			W, out, N, ncell, chunk, h, dt, L);
		HIPCHK(hipGetLastError());
		return 0;
	}

	int inhom_step(const double* W, int ncell, double L, double h, double dt, const double* Bext, double* out) override
	{
		if (ncell < 1 || ncell > maxCells) return set_err(-1, "ncell %d outside [1, %d]", ncell, maxCells);
		const size_t nphys = sizeof(double) * 4 * NN * (size_t)ncell;
		int r = Wt1.ensure(nphys);
		if (!r) r = Wt2.ensure(nphys);
		if (!r) r = WF.ensure(2 * nphys);
		if (!r) r = WF2.ensure(2 * nphys);
		if (r) return r;
		r = transport(W, ncell, L, h, 0.5 * dt, Wt1.template as<double>());
		if (!r) r = phys2fourier(Wt1.template as<double>(), 4 * ncell, WF.template as<double2>());
		if (!r) r = collision_step(WF.template as<double2>(), ncell, dt, Bext, WF2.template as<double2>());
		if (!r) r = fourier2phys(WF2.template as<double2>(), 4 * ncell, Wt2.template as<double>());
		if (!r && momentMask) r = moment_correct(Wt1.template as<double>(), Wt2.template as<double>(), ncell, L, momentMask);
		if (!r) r = transport(Wt2.template as<double>(), ncell, L, h, 0.5 * dt, out);
		lastCells = r ? 0 : ncell;
		lastDt = dt;
		lastSerial++;
		return r;
	}

	// The state the last periodic step's collision started from (after the first half transport) and
	// the collision operator there (with the magnetic term), in physical space: C = 2 (W_mid - W) / dt
	// from the step's FP64 midpoint state, so no further evaluation of the operator
	DevBuf Cf;
	int last_collision(int ncell, double* Win, double* C) override
	{
		if (lastCells < 1 || ncell != lastCells)
			return set_err(-1, "no periodic step on %d cells to take the collision operator from", ncell);
		const size_t nphys = sizeof(double) * 4 * NN * (size_t)ncell;
		if (Win) HIPCHK(hipMemcpyAsync(Win, Wt1.p, nphys, hipMemcpyDeviceToDevice, stream));
		if (!C) return 0;
		int r = Cf.ensure(2 * nphys);
		if (r) return r;
		const long long n = 4LL * NN * ncell;
		hipLaunchKernelGGL(k_scaled_diff, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
			Wmid.template as<double2>(), WF.template as<double2>(), 2.0 / lastDt, Cf.template as<double2>(), n);
		HIPCHK(hipGetLastError());
		return fourier2phys(Cf.template as<double2>(), 4 * ncell, C);
	}

	// Dirichlet step (src/simulation.c:562-830): half transport with Dirichlet ends, collision on
	// the interior cells 1 .. ncell-2 only (the end cells keep their transported values), half
	// transport.  The conservation correction applies to the collided cells.
	int inhom_step_dirichlet(const double* W, int ncell, double L, double h, double dt, const double* Bext,
		double* out) override
	{
		if (ncell < 2 || ncell > maxCells) return set_err(-1, "ncell %d outside [2, %d]", ncell, maxCells);
		const size_t cellB = sizeof(double) * 4 * NN;
		const size_t nphys = cellB * (size_t)ncell;
		int r = Wt1.ensure(nphys);
		if (!r) r = Wt2.ensure(nphys);
		if (!r) r = WF.ensure(2 * nphys);
		if (!r) r = WF2.ensure(2 * nphys);
		if (r) return r;
		double* t1 = Wt1.template as<double>();
		double* t2 = Wt2.template as<double>();
		r = transport_dirichlet(W, ncell, L, h, 0.5 * dt, t1);
		if (!r) HIPCHK(hipMemcpyAsync(t2, t1, nphys, hipMemcpyDeviceToDevice, stream));
		const int ni = ncell - 2;
		if (!r && ni > 0) {
			const size_t off = 4 * NN;   // one cell, in doubles
			r = phys2fourier(t1 + off, 4 * ni, WF.template as<double2>());
			if (!r) r = collision_step(WF.template as<double2>(), ni, dt, Bext ? Bext + 3 : nullptr, WF2.template as<double2>());
			if (!r) r = fourier2phys(WF2.template as<double2>(), 4 * ni, t2 + off);
			if (!r && momentMask) r = moment_correct(t1 + off, t2 + off, ni, L, momentMask);
		}
		if (!r) r = transport_dirichlet(t2, ncell, L, h, 0.5 * dt, out);
		return r;
	}

	// Maxwell step (src/simulation.c:882-1165): half transport with Maxwell reflection, collision of
	// every cell, half transport; the trace fluxes are taken from the state each half transport starts from.
	int inhom_step_maxwell(const double* W, int ncell, double L, double h, double dt, double lambda,
		const double* WL, const double* WR, const double* Bext, double* out) override
	{
		if (ncell < 1 || ncell > maxCells) return set_err(-1, "ncell %d outside [1, %d]", ncell, maxCells);
		const size_t nphys = sizeof(double) * 4 * NN * (size_t)ncell;
		int r = Wt1.ensure(nphys);
		if (!r) r = Wt2.ensure(nphys);
		if (!r) r = WF.ensure(2 * nphys);
		if (!r) r = WF2.ensure(2 * nphys);
		if (r) return r;
		double* t1 = Wt1.template as<double>();
		double* t2 = Wt2.template as<double>();
		r = transport_maxwell(W, ncell, L, h, 0.5 * dt, lambda, WL, WR, t1);
		if (!r) r = phys2fourier(t1, 4 * ncell, WF.template as<double2>());
		if (!r) r = collision_step(WF.template as<double2>(), ncell, dt, Bext, WF2.template as<double2>());
		if (!r) r = fourier2phys(WF2.template as<double2>(), 4 * ncell, t2);
		if (!r && momentMask) r = moment_correct(t1, t2, ncell, L, momentMask);
		if (!r) r = transport_maxwell(t2, ncell, L, h, 0.5 * dt, lambda, WL, WR, out);
		return r;
	}
};

template <typename R, int N>
int make_engine(int maxCells, qbf_engine** out, int precision)
{
	auto* e = new EngineT<R, N>();
	e->gridN = N;
	e->precision = std::is_same<R, float>::value ? QBF_FP32 : QBF_FP64;
	if (precision == QBF_BF16X3) {
		e->mfmaI3 = true;
		e->precision = QBF_BF16X3;
	}
	e->maxCells = maxCells;
	if (std::is_same<R, float>::value) {
		// the paired path deviates from the complex one by 0.15 (Nyquist content of smooth states) to
		// 0.8 (grid-scale noise) times the measured ratio; below 2e-5 that stays under the FP32 path's
		// own deviation from FP64 (2.1e-5 L1 at N = 64, J = 32, M = 16)
		e->pairMode = 1;
		e->pairTol = 2e-5;
	}
	int r = e->init();
	if (r) {
		delete e;
		return r;
	}
	*out = e;
	return 0;
}

} // namespace

extern "C" {

int qbf_engine_create(int N, int precision, int maxCells, qbf_engine** out)
{
	if (!out) return set_err(-1, "null output pointer");
	*out = nullptr;
	if (maxCells < 1) return set_err(-1, "maxCells must be >= 1");
	qbf_engine* e = nullptr;
	int r;
	if (precision == QBF_FP64 && N == 16) r = make_engine<double, 16>(maxCells, &e, precision);
	else if (precision == QBF_FP64 && N == 32) r = make_engine<double, 32>(maxCells, &e, precision);
	else if (precision == QBF_FP64 && N == 64) r = make_engine<double, 64>(maxCells, &e, precision);
	else if (precision == QBF_FP32 && N == 16) r = make_engine<float, 16>(maxCells, &e, precision);
	else if (precision == QBF_FP32 && N == 32) r = make_engine<float, 32>(maxCells, &e, precision);
	else if (precision == QBF_FP32 && N == 64) r = make_engine<float, 64>(maxCells, &e, precision);
	else if (precision == QBF_BF16X3 && N == 64) r = make_engine<float, 64>(maxCells, &e, precision);
	else if (precision == QBF_FP64 && N == 128) r = make_engine<double, 128>(maxCells, &e, precision);
	else if (precision == QBF_FP32 && N == 128) r = make_engine<float, 128>(maxCells, &e, precision);
	else return set_err(-1, "unsupported (N=%d, precision=%d)", N, precision);
	if (r) return r;
	hipStream_t s;
	if (hipStreamCreate(&s) != hipSuccess) {
		delete e;
		return set_err(-2, "hipStreamCreate failed");
	}
	e->stream = s;
	e->ownStream = true;
	// the scratch array of qbf_simulate_periodic_noalloc on this device (looked up once per device)
	r = scratch_for_current_device(&e->device, &e->scratchDev);
	if (r) {
		qbf_engine_destroy(e);
		return r;
	}
	*out = e;
	return 0;
}

void qbf_engine_destroy(qbf_engine* e)
{
	if (!e) return;
	if (e->stream) hipStreamSynchronize(e->stream);
	// graphs still alive lose their executable graph (their buffers go with the engine); the handles
	// stay valid until qbf_graph_destroy, and qbf_graph_launch on them returns an error
	for (qbf_graph* g : e->graphs) {
		if (g->stream) (void)hipStreamSynchronize(g->stream);
		g->release();
		g->e = nullptr;
	}
	e->graphs.clear();
	for (hipStream_t s : e->retired) hipStreamDestroy(s);
	if (e->ownStream && e->stream) hipStreamDestroy(e->stream);
	delete e;
}

int qbf_engine_N(const qbf_engine* e) { return e ? e->gridN : 0; }
int qbf_engine_precision(const qbf_engine* e) { return e ? e->precision : -1; }

void* qbf_get_stream(qbf_engine* e) { return e ? (void*)e->stream : nullptr; }

int qbf_set_stream(qbf_engine* e, void* s)
{
	if (!e) return set_err(-1, "null engine");
	if (e->ownStream && e->stream) {
		hipStreamSynchronize(e->stream);
		if (e->liveGraphs) e->retired.push_back(e->stream);   // a live graph replays on it
		else hipStreamDestroy(e->stream);
	}
	if (s) {
		e->stream = (hipStream_t)s;
		e->ownStream = false;
	} else {
		HIPCHK(hipStreamCreate(&e->stream));
		e->ownStream = true;
	}
	return 0;
}

int qbf_profile(qbf_engine* e, int enable)
{
	if (!e) return set_err(-1, "null engine");
	e->prof = enable != 0;
	e->nev = 0;
	e->profK = 0;
	return 0;
}

int qbf_profile_read(qbf_engine* e, double* ms_total, int* launches, double* kapps)
{
	if (!e) return set_err(-1, "null engine");
	HIPCHK(hipStreamSynchronize(e->stream));
	double tot = 0;
	for (size_t i = 0; i < e->nev; i++) {
		float ms = 0;
		HIPCHK(hipEventElapsedTime(&ms, e->evs[i].first, e->evs[i].second));
		tot += ms;
	}
	*ms_total = tot;
	*launches = (int)e->nev;
	*kapps = e->profK;
	return 0;
}

int qbf_sync(qbf_engine* e)
{
	if (!e) return set_err(-1, "null engine");
	HIPCHK(hipStreamSynchronize(e->stream));
	return 0;
}

int qbf_set_quadrature(qbf_engine* e, const qbf_quad_tables* q)
{
	if (!e || !q) return set_err(-1, "null argument");
	int r = e->set_i1(QBF_TAB_I1, q->numI1, q->I1_psiR1, q->I1_psiR2);
	if (!r) r = e->set_i1(QBF_TAB_I4, q->numI4, q->I4_psiR1, q->I4_psiR2);
	if (!r) r = e->set_i2(q->numI2, q->I2_psiR1, q->I2_psiR2);
	if (!r) r = e->set_i3(q->numI3, q->I3_psiR1, q->I3_psiR2);
	return r;
}

int qbf_set_i1_table(qbf_engine* e, int slot, unsigned num, const double* p1, const double* p2)
{
	if (!e || !p1 || !p2 || num == 0) return set_err(-1, "bad argument");
	return e->set_i1(slot, num, p1, p2);
}
int qbf_set_i2_table(qbf_engine* e, unsigned num, const double* p1, const double* p2)
{
	if (!e || !p1 || !p2 || num == 0) return set_err(-1, "bad argument");
	return e->set_i2(num, p1, p2);
}
int qbf_set_i3_table(qbf_engine* e, unsigned num, const double* p1, const double* p2)
{
	if (!e || !p1 || !p2 || num == 0) return set_err(-1, "bad argument");
	return e->set_i3(num, p1, p2);
}

int qbf_collision(qbf_engine* e, const double* dW, int ncell, int mode, double* dOut, const double* dBase,
	const double* dWcb, const double* dBext, double cf)
{
	if (!e || !dW || !dOut) return set_err(-1, "null argument");
	if (mode < 1 || mode > 3) return set_err(-1, "bad mode %d", mode);
	return e->collision((const double2*)dW, ncell, mode, (double2*)dOut, (const double2*)dBase,
		(const double2*)dWcb, dBext, cf);
}

int qbf_collision_step(qbf_engine* e, const double* dW, int ncell, double dt, const double* dBext, double* dWnext)
{
	if (!e || !dW || !dWnext) return set_err(-1, "null argument");
	return e->collision_step((const double2*)dW, ncell, dt, dBext, (double2*)dWnext);
}

int qbf_I1(qbf_engine* e, int slot, const double* f, const double* g, const double* h, double* out)
{
	if (!e || !f || !g || !h || !out) return set_err(-1, "null argument");
	if (slot != QBF_TAB_I1 && slot != QBF_TAB_I4) return set_err(-1, "bad slot");
	return e->I1(slot, (const double2*)f, (const double2*)g, (const double2*)h, (double2*)out);
}
int qbf_I2(qbf_engine* e, const double* f, const double* g, const double* h, double* out)
{
	if (!e || !f || !g || !h || !out) return set_err(-1, "null argument");
	return e->I2((const double2*)f, (const double2*)g, (const double2*)h, (double2*)out);
}
int qbf_I3(qbf_engine* e, const double* f, const double* g, const double* h, double* out)
{
	if (!e || !f || !g || !h || !out) return set_err(-1, "null argument");
	return e->I3((const double2*)f, (const double2*)g, (const double2*)h, (double2*)out);
}

int qbf_phys_to_fourier(qbf_engine* e, const double* in, int ngrid, double* out)
{
	if (!e || !in || !out || ngrid < 1) return set_err(-1, "bad argument");
	return e->phys2fourier(in, ngrid, (double2*)out);
}
int qbf_fourier_to_phys(qbf_engine* e, const double* in, int ngrid, double* out)
{
	if (!e || !in || !out || ngrid < 1) return set_err(-1, "bad argument");
	return e->fourier2phys((const double2*)in, ngrid, out);
}

int qbf_transport_periodic(qbf_engine* e, const double* W, int ncell, double L, double h, double dt, double* out)
{
	if (!e || !W || !out) return set_err(-1, "null argument");
	return e->transport(W, ncell, L, h, dt, out);
}

int qbf_inhom_step_periodic(qbf_engine* e, const double* W, int ncell, double L, double h, double dt,
	const double* Bext, double* out)
{
	if (!e || !W || !out) return set_err(-1, "null argument");
	return e->inhom_step(W, ncell, L, h, dt, Bext, out);
}

int qbf_last_collision(qbf_engine* e, int ncell, double* dWin, double* dC)
{
	if (!e) return set_err(-1, "null engine");
	return e->last_collision(ncell, dWin, dC);
}


// One Strang step (qbf_inhom_step_periodic with these buffers) as a HIP graph: the step runs once
// eagerly (every buffer sized, every combination table uploaded), then is captured on the
// engine's stream, which must not be the null stream.
int qbf_graph_capture_calls(qbf_engine* e, int (*fn)(void*), void* ctx, qbf_graph** g)
{
	if (!e || !fn || !g) return set_err(-1, "qbf_graph_capture_calls: null argument");
	*g = nullptr;
	if (!e->stream) return set_err(-1, "qbf_graph_capture_calls: the null stream cannot be captured");
	// once eagerly: sizes every buffer and uploads every combination table the calls need
	int r = fn(ctx);
	if (r) return r;
	HIPCHK(hipStreamSynchronize(e->stream));
	const bool prof = e->prof;
	e->prof = false;
	HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
	const unsigned serialBefore = e->lastSerial;
	tl_capturing = true;
	r = fn(ctx);
	tl_capturing = false;
	e->prof = prof;
	const bool setsLast = e->lastSerial != serialBefore;
	hipGraph_t graph = nullptr;
	const hipError_t ce = hipStreamEndCapture(e->stream, &graph);
	if (r || ce != hipSuccess) {
		if (graph) (void)hipGraphDestroy(graph);
		return r ? r : set_err(-2, "hipStreamEndCapture failed: %s", hipGetErrorString(ce));
	}
	auto* ng = new qbf_graph();
	ng->graph = graph;
	ng->stream = e->stream;
	ng->e = e;
	ng->setsLast = setsLast;
	ng->lastCells = e->lastCells;
	ng->lastDt = e->lastDt;
	const hipError_t ie = hipGraphInstantiate(&ng->exec, graph, nullptr, nullptr, 0);
	if (ie != hipSuccess) {
		ng->release();
		delete ng;
		return set_err(-2, "hipGraphInstantiate failed: %s", hipGetErrorString(ie));
	}
	if (hipEventCreateWithFlags(&ng->before, hipEventDisableTiming) != hipSuccess ||
		hipEventCreateWithFlags(&ng->after, hipEventDisableTiming) != hipSuccess) {
		ng->release();
		delete ng;
		return set_err(-2, "hipEventCreate failed");
	}
	e->liveGraphs++;
	e->graphs.push_back(ng);
	*g = ng;
	return 0;
}

namespace {
struct StepCall {
	qbf_engine* e;
	const double *W, *Bext;
	int ncell;
	double L, h, dt;
	double* out;
};
int step_call(void* p)
{
	auto* c = (StepCall*)p;
	return c->e->inhom_step(c->W, c->ncell, c->L, c->h, c->dt, c->Bext, c->out);
}
} // namespace

int qbf_graph_capture_inhom_step_periodic(qbf_engine* e, const double* W, int ncell, double L, double h, double dt,
	const double* Bext, double* out, qbf_graph** g)
{
	if (!e || !W || !out || !g) return set_err(-1, "qbf_graph_capture_inhom_step_periodic: null argument");
	if (e->pairMode == 1)
		return set_err(-1, "qbf_graph_capture_inhom_step_periodic: needs a fixed paired-path choice "
			"(qbf_set_real_pairs mode 0 or 2)");
	StepCall c{e, W, Bext, ncell, L, h, dt, out};
	return qbf_graph_capture_calls(e, step_call, &c, g);
}

// A replay runs on the capture stream, after the work already queued on the engine's current stream
// and before any work queued there later (the graph shares the engine's buffers with eager calls).
int qbf_graph_launch(qbf_graph* g)
{
	if (!g) return set_err(-1, "null graph");
	if (!g->e || !g->exec) return set_err(-1, "qbf_graph_launch: the engine of this graph was destroyed");
	hipStream_t cur = g->e->stream;
	const bool other = cur != g->stream;
	if (other) {
		HIPCHK(hipEventRecord(g->before, cur));
		HIPCHK(hipStreamWaitEvent(g->stream, g->before, 0));
	}
	HIPCHK(hipGraphLaunch(g->exec, g->stream));
	if (other) {
		HIPCHK(hipEventRecord(g->after, g->stream));
		HIPCHK(hipStreamWaitEvent(cur, g->after, 0));
	}
	if (g->setsLast) {
		g->e->lastCells = g->lastCells;
		g->e->lastDt = g->lastDt;
	}
	return 0;
}

void qbf_graph_destroy(qbf_graph* g)
{
	if (!g) return;
	if (g->e) {
		if (g->stream) (void)hipStreamSynchronize(g->stream);
		auto& v = g->e->graphs;
		for (size_t i = 0; i < v.size(); i++)
			if (v[i] == g) {
				v.erase(v.begin() + (long)i);
				break;
			}
		g->e->liveGraphs--;
	}
	g->release();
	delete g;
}

int qbf_inhom_step_dirichlet(qbf_engine* e, const double* W, int ncell, double L, double h, double dt,
	const double* Bext, double* out)
{
	if (!e || !W || !out) return set_err(-1, "null argument");
	return e->inhom_step_dirichlet(W, ncell, L, h, dt, Bext, out);
}

int qbf_transport_dirichlet(qbf_engine* e, const double* W, int ncell, double L, double h, double dt, double* out)
{
	if (!e || !W || !out) return set_err(-1, "null argument");
	return e->transport_dirichlet(W, ncell, L, h, dt, out);
}

int qbf_inhom_step_maxwell(qbf_engine* e, const double* W, int ncell, double L, double h, double dt, double lambda,
	const double* WMaxwL, const double* WMaxwR, const double* Bext, double* out)
{
	if (!e || !W || !WMaxwL || !WMaxwR || !out) return set_err(-1, "qbf_inhom_step_maxwell: null argument");
	return e->inhom_step_maxwell(W, ncell, L, h, dt, lambda, WMaxwL, WMaxwR, Bext, out);
}

int qbf_transport_maxwell(qbf_engine* e, const double* W, int ncell, double L, double h, double dt, double lambda,
	const double* WMaxwL, const double* WMaxwR, double* out)
{
	if (!e || !W || !WMaxwL || !WMaxwR || !out) return set_err(-1, "qbf_transport_maxwell: null argument");
	return e->transport_maxwell(W, ncell, L, h, dt, lambda, WMaxwL, WMaxwR, out);
}

int qbf_set_moment_correction(qbf_engine* e, int mask)
{
	if (!e || mask < 0 || mask > 7) return set_err(-1, "bad argument");
	e->momentMask = mask;
	return 0;
}

// boundary of a device-resident simulation loop: periodic, Dirichlet or Maxwell (lambda and the two
// boundary states, device)
struct Boundary {
	int kind = 0;                     // 0 periodic, 1 Dirichlet, 2 Maxwell
	double lambda = 0.0;
	const double* WL = nullptr;
	const double* WR = nullptr;
};

// the device-resident loop of qbf_simulate_{periodic,dirichlet,maxwell} over two state buffers
static int simulate_loop(qbf_engine* e, const double* dW0, int ncell, double L, double h, double dt,
	unsigned numsteps, unsigned stride, const double* dBext, double* hostFrames, double* const buf[2], size_t bytes,
	const Boundary& bc = Boundary())
{
	int r = 0;
	const double* cur = dW0;
	size_t f = 0;
	// frame 0 is the initial state (src/simulation.c:1294)
	if (hipMemcpyAsync(hostFrames, dW0, bytes, hipMemcpyDeviceToHost, e->stream) != hipSuccess)
		r = set_err(-2, "qbf_simulate: copy of frame 0 failed");
	f = 1;
	for (unsigned it = 1; it < numsteps && !r; it++) {
		double* nxt = buf[it & 1];
		if (bc.kind == 1)
			r = e->inhom_step_dirichlet(cur, ncell, L, h, dt, dBext, nxt);
		else if (bc.kind == 2)
			r = e->inhom_step_maxwell(cur, ncell, L, h, dt, bc.lambda, bc.WL, bc.WR, dBext, nxt);
		else
			r = e->inhom_step(cur, ncell, L, h, dt, dBext, nxt);
		cur = nxt;
		if (!r && it % stride == 0) {
			if (hipMemcpyAsync((char*)hostFrames + f * bytes, cur, bytes, hipMemcpyDeviceToHost, e->stream) != hipSuccess)
				r = set_err(-2, "qbf_simulate: frame copy failed");
			f++;
		}
	}
	if (hipStreamSynchronize(e->stream) != hipSuccess && !r) r = set_err(-2, "qbf_simulate: stream error");
	return r;
}

// two state buffers of hipMalloc'ed device memory around simulate_loop
static int simulate_alloc(qbf_engine* e, const double* dW0, int ncell, double L, double h, double dt,
	unsigned numsteps, unsigned stride, const double* dBext, double* hostFrames, const Boundary& bc, const char* who)
{
	if (!e || !dW0 || !hostFrames || ncell < 1 || stride < 1) return set_err(-1, "%s: bad argument", who);
	const size_t bytes = sizeof(double) * 4 * (size_t)e->gridN * e->gridN * ncell;
	double* buf[2] = {nullptr, nullptr};
	if (hipMalloc(&buf[0], bytes) != hipSuccess || hipMalloc(&buf[1], bytes) != hipSuccess) {
		if (buf[0]) (void)hipFree(buf[0]);
		return set_err(-3, "%s: hipMalloc(%zu) failed", who, bytes);
	}
	const int r = simulate_loop(e, dW0, ncell, L, h, dt, numsteps, stride, dBext, hostFrames, buf, bytes, bc);
	(void)hipFree(buf[0]);
	(void)hipFree(buf[1]);
	return r;
}

int qbf_simulate_periodic(qbf_engine* e, const double* dW0, int ncell, double L, double h, double dt,
	unsigned numsteps, unsigned stride, const double* dBext, double* hostFrames)
{
	return simulate_alloc(e, dW0, ncell, L, h, dt, numsteps, stride, dBext, hostFrames, Boundary(),
		"qbf_simulate_periodic");
}

int qbf_simulate_dirichlet(qbf_engine* e, const double* dW0, int ncell, double L, double h, double dt,
	unsigned numsteps, unsigned stride, const double* dBext, double* hostFrames)
{
	Boundary bc;
	bc.kind = 1;
	return simulate_alloc(e, dW0, ncell, L, h, dt, numsteps, stride, dBext, hostFrames, bc, "qbf_simulate_dirichlet");
}

int qbf_simulate_maxwell(qbf_engine* e, const double* dW0, int ncell, double L, double h, double dt, double lambda,
	const double* dWMaxwL, const double* dWMaxwR, unsigned numsteps, unsigned stride, const double* dBext,
	double* hostFrames)
{
	if (!dWMaxwL || !dWMaxwR) return set_err(-1, "qbf_simulate_maxwell: null boundary state");
	Boundary bc;
	bc.kind = 2;
	bc.lambda = lambda;
	bc.WL = dWMaxwL;
	bc.WR = dWMaxwR;
	return simulate_alloc(e, dW0, ncell, L, h, dt, numsteps, stride, dBext, hostFrames, bc, "qbf_simulate_maxwell");
}

int qbf_simulate_periodic_noalloc(qbf_engine* e, const double* dW0, int ncell, double L, double h, double dt,
	unsigned numsteps, unsigned stride, const double* dBext, double* hostFrames)
{
	if (!e || !dW0 || !hostFrames || ncell < 1 || stride < 1)
		return set_err(-1, "qbf_simulate_periodic_noalloc: bad argument");
	const size_t bytes = sizeof(double) * 4 * (size_t)e->gridN * e->gridN * ncell;
	if (bytes > QBF_NOALLOC_STATE_BYTES)
		return set_err(-1, "qbf_simulate_periodic_noalloc: a state of %d cells at N = %d is %zu bytes, above the "
			"limit of %zu (QBF_NOALLOC_STATE_BYTES)", ncell, e->gridN, bytes, (size_t)QBF_NOALLOC_STATE_BYTES);
	if (!e->scratchDev) return set_err(-2, "qbf_simulate_periodic_noalloc: no scratch address for the engine's device");
	int dev = -1;
	if (hipGetDevice(&dev) != hipSuccess || dev != e->device)
		return set_err(-1, "qbf_simulate_periodic_noalloc: the current device (%d) is not the engine's (%d)", dev,
			e->device);
	// the two state buffers: both halves of the device's scratch array (every element is written
	// by a step before it is read, so nothing needs clearing)
	double* const buf[2] = {(double*)e->scratchDev, (double*)((char*)e->scratchDev + QBF_NOALLOC_STATE_BYTES)};
	tl_noalloc = true;
	const int r = simulate_loop(e, dW0, ncell, L, h, dt, numsteps, stride, dBext, hostFrames, buf, bytes);
	tl_noalloc = false;
	return r;
}

int qbf_cell_diagnostics(qbf_engine* e, const double* dW, int ncell, double L, double* dOut)
{
	if (!e || !dW || !dOut || ncell < 1) return set_err(-1, "qbf_cell_diagnostics: bad argument");
	hipLaunchKernelGGL(k_cell_diag, dim3(ncell), dim3(256), 0, e->stream, dW, e->gridN, L, dOut);
	HIPCHK(hipGetLastError());
	return 0;
}

int qbf_set_real_pairs(qbf_engine* e, int mode, double tol)
{
	if (!e || mode < 0 || mode > 2 || !(tol >= 0)) return set_err(-1, "qbf_set_real_pairs: bad argument");
	e->pairMode = mode;
	e->pairTol = tol;
	return 0;
}

int qbf_real_pairs_status(qbf_engine* e, int* used, double* ratio)
{
	if (!e) return set_err(-1, "qbf_real_pairs_status: null engine");
	if (used) *used = e->pairUsed;
	if (ratio) *ratio = e->pairRatio;
	return 0;
}

int qbf_fermi_dirac(qbf_engine* e, const double* dW, int ncell, double L, double* dParams, double* dWFD)
{
	if (!e || !dW || !dParams || ncell < 1) return set_err(-1, "qbf_fermi_dirac: bad argument");
	hipLaunchKernelGGL(k_fermi_dirac, dim3(ncell), dim3(256), 0, e->stream, dW, e->gridN, L, dParams, dWFD);
	HIPCHK(hipGetLastError());
	return 0;
}

int qbf_fd_distance(qbf_engine* e, const double* dW, int ncell, double L, const double* dWFD, int fd_stride,
	double* dOut)
{
	if (!e || !dW || !dWFD || !dOut || ncell < 1 || (fd_stride != 0 && fd_stride != 1))
		return set_err(-1, "qbf_fd_distance: bad argument");
	hipLaunchKernelGGL(k_fd_distance, dim3(ncell), dim3(256), 0, e->stream, dW, dWFD, fd_stride, e->gridN, L, dOut);
	HIPCHK(hipGetLastError());
	return 0;
}

int qbf_cell_average(qbf_engine* e, const double* dW, int ncell, double* dOut)
{
	if (!e || !dW || !dOut || ncell < 1) return set_err(-1, "qbf_cell_average: bad argument");
	const long long nel = 4LL * e->gridN * e->gridN;
	hipLaunchKernelGGL(k_cell_average, dim3((unsigned)((nel + 255) / 256)), dim3(256), 0, e->stream, dW, ncell, nel, dOut);
	HIPCHK(hipGetLastError());
	return 0;
}

int qbf_moment_correct(qbf_engine* e, const double* Wold, double* Wnew, int ncell, double L, int mask)
{
	if (!e || !Wold || !Wnew || ncell < 1 || mask < 0 || mask > 7) return set_err(-1, "bad argument");
	if (e->moment_correct(Wold, Wnew, ncell, L, mask)) return set_err(-2, "k_moment_correct launch failed");
	return 0;
}

int qbf_transport_line_periodic(qbf_engine* e, const double* U, int n, double A, double h, double dt, double* out)
{
	if (!e || !U || !out || n < 1) return set_err(-1, "bad argument");
	hipLaunchKernelGGL(k_transport_line, dim3((n + 255) / 256), dim3(256), 0, e->stream, U, out, n, A, h, dt);
	HIPCHK(hipGetLastError());
	return 0;
}

void* qbf_malloc(size_t bytes)
{
	void* p = nullptr;
	if (hipMalloc(&p, bytes) != hipSuccess) {
		set_err(-3, "hipMalloc(%zu) failed", bytes);
		return nullptr;
	}
	return p;
}

void qbf_free(void* p)
{
	if (p) hipFree(p);
}

int qbf_h2d(qbf_engine* e, void* dst, const void* src, size_t bytes)
{
	HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, e ? e->stream : nullptr));
	HIPCHK(hipStreamSynchronize(e ? e->stream : nullptr));
	return 0;
}

int qbf_d2h(qbf_engine* e, void* dst, const void* src, size_t bytes)
{
	HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, e ? e->stream : nullptr));
	HIPCHK(hipStreamSynchronize(e ? e->stream : nullptr));
	return 0;
}

} // extern "C"
