// qbf_kernels.hip -- HIP kernels of the collision operator and the transport step (gfx950).
//
// All collision kernels work on zero-padded P x P grids (P = 2N) held by one
// workgroup (see qbf_device.hpp for the register layouts).  The collision
// operator C = Cc + Cd of the reference (src/collision.c:74-168, 209-254) is
// evaluated in an exactly reassociated form: every term that the reference
// convolves, pads, extracts and sums separately is summed in the transformed
// domain instead, so each Cc+Cd evaluation needs 15 + 22J + 36JM padded FFTs
// instead of the reference's 68 + 96J + 24J^2 + 80JM.
#include "qbf_device.hpp"
#include "qbf_kernels.hpp"

namespace qbf {

template <int N>
__device__ __forceinline__ int unpad(int x)
{
	// padded frequency index (axis [0..N-1, -N..-1]) -> index of the N x N grid
	// (axis [0..N/2-1, -N/2..-1]) or -1 for a padding zero (src/grid.c:171-188)
	return x < N / 2 ? x : (x >= 2 * N - N / 2 ? x - N : -1);
}

// ---------------------------------------------------------------------------------------------
// pruned FFT2 of pad(pre .* src): C_i load -> column FFT -> LDS -> row FFT -> R_b (registers)
// The padded index p + G s of element s of a lane is a padding zero exactly when
// E/4 <= s < 3E/4 (P = 2N = G E, lane p < G), the same for every lane: those inputs are
// compile-time zeros in both passes, the zero columns are neither written to nor read from LDS.
template <int E>
__device__ __forceinline__ constexpr bool padded_nonzero(int s) { return s < E / 4 || s >= 3 * E / 4; }

template <typename R, int N, class GE>
__device__ __forceinline__ void padfft_reg(cx<R>* v, const double2* __restrict__ src, const void* pre, int preKind,
	const Xpose<R, GE>& X, const cx<R>* tw, int line, int p)
{
	constexpr int G = GE::G, E = GE::E, P = GE::P;
	static_assert(2 * N == G * E, "padded line of 2N = G E elements");
	const int kn = unpad<N>(line);
	// base pointers formed here (opaque SGPRs): in a loop over terms the per-element 64-bit
	// addresses are then not hoisted into (spilled) registers
	const double2* sp = uni(src, 0);
	const double* p1 = uni((const double*)pre, 0);
	const double2* p2 = uni((const double2*)pre, 0);
	if (kn >= 0) {
#pragma unroll
		for (int s = 0; s < E; s++) {
			if (padded_nonzero<E>(s)) {
				const int ln = p + G * s - (s >= E / 4 ? N : 0);
				const int e = ln * N + kn;
				double2 x = sp[e];
				if (preKind == 1) {
					const double r = p1[e];
					x.x *= r;
					x.y *= r;
				} else if (preKind == 2) {
					const double2 c = p2[e];
					x = make_double2(c.x * x.x - c.y * x.y, c.x * x.y + c.y * x.x);
				}
				v[s] = mk<R>((R)x.x, (R)x.y);
			} else {
				v[s] = czero<R>();
			}
		}
		line_fwd<R, P, G, E>(v, p, tw);
	}
	__syncthreads();
	if (kn >= 0) X.put_Cb(v, line, p);
	__syncthreads();
#pragma unroll
	for (int s = 0; s < E; s++) v[s] = padded_nonzero<E>(s) ? X.get(line, p + G * s) : czero<R>();
	line_fwd<R, P, G, E>(v, p, tw);
}

// R_b (physical) -> IFFT2 -> C_i (frequency)
template <typename R, class GE>
__device__ __forceinline__ void ifft2_reg(cx<R>* v, const Xpose<R, GE>& X, const cx<R>* tw, int line, int p)
{
	constexpr int G = GE::G, E = GE::E, P = GE::P;
	line_bwd<R, P, G, E>(v, p, tw);
	__syncthreads();
	X.put_Ri(v, line, p);
	__syncthreads();
	X.get_Cb(v, line, p);
	line_bwd<R, P, G, E>(v, p, tw);
}

// C_i (frequency) -> FFT2 -> R_b (physical)
template <typename R, class GE>
__device__ __forceinline__ void fft2_reg(cx<R>* v, const Xpose<R, GE>& X, const cx<R>* tw, int line, int p)
{
	constexpr int G = GE::G, E = GE::E, P = GE::P;
	line_fwd<R, P, G, E>(v, p, tw);
	__syncthreads();
	X.put_Cb(v, line, p);
	__syncthreads();
	X.get_Ri(v, line, p);
	line_fwd<R, P, G, E>(v, p, tw);
}

// ---------------------------------------------------------------------------------------------
template <typename R, int N, int T, bool XG>
__global__ __launch_bounds__(T) void k_padfft(PadFFTArgs a)
{
	// Two grids per workgroup: the N non-zero columns of grid 2b on lines 0..N-1 and those of grid
	// 2b+1 on lines N..2N-1 (every lane works in the column pass); their transposed columns fill
	// the LDS grid, columns 0..N-1 and N..2N-1; then a row pass per grid.
	using GE = Geo<2 * N, T>;
	constexpr int P = GE::P, E = GE::E, G = GE::G;
	__shared__ cx<R> lds[XG ? 1 : Xpose<R, GE>::SIZE];
	__shared__ cx<R> tw[TwTab<P, T>::SIZE];
	load_twiddles<R, P, T>(tw, a.tw256);
	cx<R>* xbuf = XG ? (cx<R>*)a.xs + (size_t)blockIdx.x * Xpose<R, GE>::SIZE : lds;
	const int o0 = 2 * blockIdx.x;
	const int line = threadIdx.x / G, p = threadIdx.x % G;
	const int o = o0 + line / N, kn = line % N;
	Xpose<R, GE> X{xbuf};
	cx<R> v[E];
	__syncthreads();
	if (o < a.ngrid) {
		const int slot = o % a.nslot;
		const int qq = (o / a.nslot) % a.nq;
		const int cell = o / (a.nslot * a.nq);
		const double2* src = a.src + ((size_t)cell * a.nsrc + a.srcOfSlot[slot]) * (N * N);
		const size_t poff = (size_t)(a.q0 + qq) * (N * N);
#pragma unroll
		for (int s = 0; s < E; s++) {
			if (padded_nonzero<E>(s)) {
				const int e = (p + G * s - (s >= E / 4 ? N : 0)) * N + kn;
				double2 x = src[e];
				if (a.preKind == 1) {
					const double r = ((const double*)a.pre)[poff + e];
					x.x *= r;
					x.y *= r;
				} else if (a.preKind == 2) {
					const double2 c = ((const double2*)a.pre)[poff + e];
					x = make_double2(c.x * x.x - c.y * x.y, c.x * x.y + c.y * x.x);
				}
				v[s] = mk<R>((R)x.x, (R)x.y);
			} else {
				v[s] = czero<R>();
			}
		}
		line_fwd<R, P, G, E>(v, p, tw);
		X.put_Cb(v, line, p);
	}
	__syncthreads();
	for (int g = 0; g < 2 && o0 + g < a.ngrid; g++) {
#pragma unroll
		for (int s = 0; s < E; s++)
			v[s] = padded_nonzero<E>(s) ? X.get(line, g * N + p + G * s - (s >= E / 4 ? N : 0)) : czero<R>();
		line_fwd<R, P, G, E>(v, p, tw);
		store_native<R, GE>((cx<R>*)a.out + (size_t)(o0 + g) * (P * P), v);
	}
}

// ---------------------------------------------------------------------------------------------
template <typename R, int N, int T, bool XG>
__global__ __launch_bounds__(T) void k_i1acc(I1Args a)
{
	using GE = Geo<2 * N, T>;
	constexpr int P = GE::P, E = GE::E;
	__shared__ cx<R> lds[XG ? 1 : Xpose<R, GE>::SIZE];
	__shared__ cx<R> tw[TwTab<P, T>::SIZE];
	load_twiddles<R, P, T>(tw, a.tw256);
	cx<R>* xbuf = XG ? (cx<R>*)a.xs + (size_t)blockIdx.x * Xpose<R, GE>::SIZE : lds;
	const int lb = xcd_block();
	const int chunk = lb % a.nchunk;
	const int inst = (lb / a.nchunk) % a.ninst;
	const int cell = lb / (a.nchunk * a.ninst);
	const int qa = (int)(((long long)chunk * a.nterm) / a.nchunk);
	const int qb = (int)(((long long)(chunk + 1) * a.nterm) / a.nchunk);
	const int line = threadIdx.x / GE::G, p = threadIdx.x % GE::G;
	const double2* g = a.src + ((size_t)cell * a.nsrc + a.instSrc[inst]) * (N * N);
	const cx<R>* Fh = (const cx<R>*)a.Fhat + ((size_t)cell * a.nFh + a.instFh[inst]) * (P * P);
	const R fa = (R)a.fa, fb = (R)a.fb;
	Xpose<R, GE> X{xbuf};
	cx<R> acc[E], v[E];
#pragma unroll
	for (int s = 0; s < E; s++) acc[s] = czero<R>();
	__syncthreads();
	for (int q = qa; q < qb; q++) {
		padfft_reg<R, N, GE>(v, g, a.phi + (size_t)q * (N * N), 1, X, tw, line, p);
#pragma unroll
		for (int u = 0; u < E; u++) {
			cx<R> f = uni(Fh, (long long)u * T)[threadIdx.x];
			f = mk<R>(fa * f.x + fb, fa * f.y);
			v[u] = cmul<R>(f, v[u]);
		}
		ifft2_reg<R, GE>(v, X, tw, line, p);
		const R* ps = (const R*)a.psi1n + (size_t)q * (P * P);
#pragma unroll
		for (int s = 0; s < E; s++) {
			const R w = uni(ps, (long long)s * T)[threadIdx.x];
			acc[s].x += w * v[s].x;
			acc[s].y += w * v[s].y;
		}
	}
	fft2_reg<R, GE>(acc, X, tw, line, p);
	store_native<R, GE>((cx<R>*)a.out + (((size_t)cell * a.ninst + inst) * a.nchunk + chunk) * (P * P), acc);
}

// ---------------------------------------------------------------------------------------------
// PAIR (f = g, real grids): FFT2(pad(psi1 f)) and FFT2(pad(psi2 f)) are the real and imaginary
// parts of FFT2(pad((psi1 + i psi2) f)) when both are real (psi1, psi2 even, f the spectrum of a
// real state with empty Nyquist lines), so one transform per term gives the product re * im.
template <typename R, int N, int T, bool XG, bool PAIR>
__global__ __launch_bounds__(T) void k_i2acc(I2Args a)
{
	using GE = Geo<2 * N, T>;
	constexpr int P = GE::P, E = GE::E;
	__shared__ cx<R> lds[XG ? 1 : Xpose<R, GE>::SIZE];
	__shared__ cx<R> tw[TwTab<P, T>::SIZE];
	load_twiddles<R, P, T>(tw, a.tw256);
	cx<R>* xbuf = XG ? (cx<R>*)a.xs + (size_t)blockIdx.x * Xpose<R, GE>::SIZE : lds;
	const int lb = xcd_block();
	const int chunk = lb % a.nchunk;
	const int inst = (lb / a.nchunk) % a.ninst;
	const int cell = lb / (a.nchunk * a.ninst);
	const int qa = (int)(((long long)chunk * a.nterm) / a.nchunk);
	const int qb = (int)(((long long)(chunk + 1) * a.nterm) / a.nchunk);
	const int line = threadIdx.x / GE::G, p = threadIdx.x % GE::G;
	const double2* f = a.src + ((size_t)cell * a.nsrc + a.instF[inst]) * (N * N);
	const double2* g = a.src + ((size_t)cell * a.nsrc + a.instG[inst]) * (N * N);
	Xpose<R, GE> X{xbuf};
	cx<R> acc[E], x1[E], v[E];
#pragma unroll
	for (int u = 0; u < E; u++) acc[u] = czero<R>();
	__syncthreads();
	for (int q = qa; q < qb; q++) {
		if constexpr (PAIR) {
			padfft_reg<R, N, GE>(v, f, a.psic + (size_t)q * (N * N), 2, X, tw, line, p);
#pragma unroll
			for (int u = 0; u < E; u++) acc[u].x += v[u].x * v[u].y;
		} else {
			padfft_reg<R, N, GE>(x1, f, a.psi1 + (size_t)q * (N * N), 1, X, tw, line, p);
			padfft_reg<R, N, GE>(v, g, a.psi2 + (size_t)q * (N * N), 1, X, tw, line, p);
#pragma unroll
			for (int u = 0; u < E; u++) {
				const cx<R> t = cmul<R>(x1[u], v[u]);
				acc[u] = cadd<R>(acc[u], t);
			}
		}
	}
	store_native<R, GE>((cx<R>*)a.out + (((size_t)cell * a.ninst + inst) * a.nchunk + chunk) * (P * P), acc);
}

// ---------------------------------------------------------------------------------------------
// The dominant kernel: 16 applications of K_q per quadrature term and cell for CdInt.
// PAIR: the grids hold packed pairs of real grids (real part: component 2p, imaginary part:
// component 2p+1), S and F from the packed spectra W_2p + i W_2p+1 and H one component of a packed
// F grid (instH = 2 slot + component).  K_q maps real grids to real grids, so one complex K_q
// application carries the two instances (S_2p H, S_2p+1 H) and acc += w_2p F_2p Y_2p + w_2p+1 F_2p+1 Y_2p+1
// (instW[out][2 in], [2 in + 1]).  Exact for spectra of real states whose Nyquist lines vanish;
// otherwise off by the dropped imaginary parts, which the host bounds before choosing it.
template <typename R, int N, int T, bool XG, bool PAIR>
__global__ __launch_bounds__(T) void k_i3acc(I3Args a)
{
	using GE = Geo<2 * N, T>;
	constexpr int P = GE::P, E = GE::E;
	__shared__ cx<R> lds[XG ? 1 : Xpose<R, GE>::SIZE];
	__shared__ cx<R> tw[TwTab<P, T>::SIZE];
	load_twiddles<R, P, T>(tw, a.tw256);
	cx<R>* xbuf = XG ? (cx<R>*)a.xs + (size_t)blockIdx.x * Xpose<R, GE>::SIZE : lds;
	const int lb = xcd_block();
	const int chunk = lb % a.nchunk;
	const int out = (lb / a.nchunk) % a.nout;
	const int cell = lb / (a.nchunk * a.nout);
	const int line = threadIdx.x / GE::G, p = threadIdx.x % GE::G;
	const int t = threadIdx.x;
	const int qa = (int)(((long long)chunk * a.nq) / a.nchunk);
	const int qb = (int)(((long long)(chunk + 1) * a.nq) / a.nchunk);
	cx<R>* accp = (cx<R>*)a.acc + (((size_t)cell * a.nout + out) * a.nchunk + chunk) * (P * P);
	Xpose<R, GE> X{xbuf};
	cx<R> acc[E], v[E];
	if (a.accumulate) {
		load_native<R, GE>(acc, accp);
	} else {
#pragma unroll
		for (int u = 0; u < E; u++) acc[u] = czero<R>();
	}
	const cx<R>* Sh = (const cx<R>*)a.Shat + (size_t)cell * a.nS * (P * P);
	const int ni = a.ninst[out];
	__syncthreads();
	for (int qq = qa; qq < qb; qq++) {
		const cx<R>* Fq = (const cx<R>*)a.F + ((size_t)cell * a.nq + qq) * a.nF * (P * P);
		const R* ps = (const R*)a.psi2n + (size_t)(a.q0 + qq) * (P * P);
		for (int in = 0; in < ni; in++) {
			const cx<R>* Sg = Sh + (size_t)a.instS[out][in] * (P * P);
			const int hs = a.instH[out][in];
			const cx<R>* Fh = Fq + (size_t)(PAIR ? hs >> 1 : hs) * (P * P);
			const cx<R>* Ff = Fq + (size_t)a.instF[out][in] * (P * P);
			if constexpr (PAIR) {
				const bool hy = hs & 1;
#pragma unroll
				for (int u = 0; u < E; u++) {
					const cx<R> h = uni(Fh, (long long)u * T)[t];
					const R hv = hy ? h.y : h.x;
					const cx<R> sv = uni(Sg, (long long)u * T)[t];
					v[u] = mk<R>(sv.x * hv, sv.y * hv);
				}
			} else {
#pragma unroll
				for (int u = 0; u < E; u++) v[u] = cmul<R>(uni(Sg, (long long)u * T)[t], uni(Fh, (long long)u * T)[t]);
			}
			ifft2_reg<R, GE>(v, X, tw, line, p);
			// keep the table and operand loads next to their use: hoisting them above the
			// line FFTs would keep E extra complex values live and spill the accumulator
			__builtin_amdgcn_sched_barrier(0);
#pragma unroll
			for (int s = 0; s < E; s++) {
				const R m = uni(ps, (long long)s * T)[t];
				v[s].x *= m;
				v[s].y *= m;
			}
			fft2_reg<R, GE>(v, X, tw, line, p);
			__builtin_amdgcn_sched_barrier(0);
			if constexpr (PAIR) {
				const R wa = (R)a.instW[out][2 * in], wb = (R)a.instW[out][2 * in + 1];
#pragma unroll
				for (int u = 0; u < E; u++) {
					const cx<R> f = uni(Ff, (long long)u * T)[t];
					acc[u].x += (wa * f.x) * v[u].x + (wb * f.y) * v[u].y;
				}
			} else {
				const R w = (R)a.instW[out][in];
#pragma unroll
				for (int u = 0; u < E; u++) {
					const cx<R> y = cmul<R>(uni(Ff, (long long)u * T)[t], v[u]);
					acc[u].x += w * y.x;
					acc[u].y += w * y.y;
				}
			}
			__builtin_amdgcn_sched_barrier(0);
		}
	}
	store_native<R, GE>(accp, acc);
}

// k_padfft with each workgroup looping over pairs of grids (pair b, b + gridDim.x, ...): the state
// values of its next pair are loaded before the current pair's row passes, so that their latency
// overlaps the row FFTs and the stores, and one resident workgroup per CU (the LDS exchange takes
// 141.5 KiB) runs from the first pair to its last.  Same arithmetic in the same order as k_padfft
// (LDS exchange only).
template <typename R, int N, int T>
__global__ __launch_bounds__(T) void k_padfft_loop(PadFFTArgs a)
{
	using GE = Geo<2 * N, T>;
	constexpr int P = GE::P, E = GE::E, G = GE::G;
	constexpr int NZ = E / 2;                      // non-zero elements per lane in the column pass
	__shared__ cx<R> lds[Xpose<R, GE>::SIZE];
	__shared__ cx<R> tw[TwTab<P, T>::SIZE];
	load_twiddles<R, P, T>(tw, a.tw256);
	const int line = threadIdx.x / G, p = threadIdx.x % G;
	const int kn = line % N;
	const int npair = (a.ngrid + 1) / 2;
	Xpose<R, GE> X{lds};
	cx<R> v[E];
	double2 raw[NZ];
	// the state values of pair b for this lane (the grid o = 2 b + line / N, if it exists)
	auto load = [&](int b) {
		const int o = 2 * b + line / N;
		if (o < a.ngrid) {
			const int slot = o % a.nslot;
			const int cell = o / (a.nslot * a.nq);
			const double2* src = a.src + ((size_t)cell * a.nsrc + a.srcOfSlot[slot]) * (N * N);
#pragma unroll
			for (int j = 0; j < NZ; j++) {
				const int s = j < NZ / 2 ? j : j + NZ;
				raw[j] = src[(p + G * s - (s >= E / 4 ? N : 0)) * N + kn];
			}
		}
	};
	int b = blockIdx.x;
	load(b);
	for (; b < npair; b += gridDim.x) {
		const int o0 = 2 * b;
		const int o = o0 + line / N;
		__syncthreads();                           // the previous pair's row passes are done with X
		if (o < a.ngrid) {
			const int qq = (o / a.nslot) % a.nq;
			const size_t poff = (size_t)(a.q0 + qq) * (N * N);
#pragma unroll
			for (int s = 0; s < E; s++) {
				if (padded_nonzero<E>(s)) {
					const int j = s < E / 4 ? s : s - NZ;
					const int e = (p + G * s - (s >= E / 4 ? N : 0)) * N + kn;
					double2 x = raw[j];
					if (a.preKind == 1) {
						const double r = ((const double*)a.pre)[poff + e];
						x.x *= r;
						x.y *= r;
					} else if (a.preKind == 2) {
						const double2 c = ((const double2*)a.pre)[poff + e];
						x = make_double2(c.x * x.x - c.y * x.y, c.x * x.y + c.y * x.x);
					}
					v[s] = mk<R>((R)x.x, (R)x.y);
				} else {
					v[s] = czero<R>();
				}
			}
			line_fwd<R, P, G, E>(v, p, tw);
			X.put_Cb(v, line, p);
		}
		__syncthreads();
		if (b + (int)gridDim.x < npair) load(b + gridDim.x);
		for (int g = 0; g < 2 && o0 + g < a.ngrid; g++) {
#pragma unroll
			for (int s = 0; s < E; s++)
				v[s] = padded_nonzero<E>(s) ? X.get(line, g * N + p + G * s - (s >= E / 4 ? N : 0)) : czero<R>();
			line_fwd<R, P, G, E>(v, p, tw);
			// non-temporal stores: the F grids of a launch (GiBs) are read back from HBM by the I3
			// kernel, so they need not displace other lines from L2 (FP32 step 329 -> 327 ms)
			cx<R>* dst = (cx<R>*)a.out + (size_t)(o0 + g) * (P * P);
#pragma unroll
			for (int u = 0; u < E; u++) {
				cx<R>* q = uni(dst, (long long)u * T) + threadIdx.x;
				__builtin_nontemporal_store(v[u].x, &q->x);
				__builtin_nontemporal_store(v[u].y, &q->y);
			}
		}
	}
}

// ---------------------------------------------------------------------------------------------
template <typename R, int N, int T, bool XG>
__global__ __launch_bounds__(T) void k_combine(CombineArgs a)
{
	using GE = Geo<2 * N, T>;
	constexpr int P = GE::P, E = GE::E, G = GE::G;
	__shared__ cx<R> lds[XG ? 1 : Xpose<R, GE>::SIZE];
	__shared__ cx<R> tw[TwTab<P, T>::SIZE];
	load_twiddles<R, P, T>(tw, a.tw256);
	cx<R>* xbuf = XG ? (cx<R>*)a.xs + (size_t)blockIdx.x * Xpose<R, GE>::SIZE : lds;
	const int o = blockIdx.x % a.nout;
	const int cell = blockIdx.x / a.nout;
	const int line = threadIdx.x / G, p = threadIdx.x % G;
	const int t = threadIdx.x;
	Xpose<R, GE> X{xbuf};
	cx<R> v[E];
#pragma unroll
	for (int u = 0; u < E; u++) v[u] = czero<R>();
	for (int it = 0; it < a.nterm[o]; it++) {
		const CTerm c = a.terms[o * a.maxTerms + it];
		const cx<R>* Xp = (const cx<R>*)c.X + (size_t)cell * c.xcell;
		const cx<R>* Yp = (const cx<R>*)c.Y + (size_t)cell * c.ycell;
		const R ax = (R)c.ax, bx = (R)c.bx, cf = (R)c.coef;
#pragma unroll
		for (int u = 0; u < E; u++) {
			cx<R> y = mk<R>(R(1), R(0));
			if (c.Y) {
				y = Yp[u * T + t];
				for (int n = 1; n < c.nsum; n++) y = cadd<R>(y, Yp[(size_t)n * c.ysum + u * T + t]);
			}
			if (c.X) {
				cx<R> x = Xp[u * T + t];
				x = mk<R>(ax * x.x + bx, ax * x.y);
				y = cmul<R>(x, y);
			}
			v[u].x += cf * y.x;
			v[u].y += cf * y.y;
		}
	}
	__syncthreads();
	ifft2_reg<R, GE>(v, X, tw, line, p);
	// C_i: this thread holds (l = p + G*s, k = line); keep the N x N block (src/grid.c:195-210)
	const int kn = unpad<N>(line);
	if (kn < 0) return;
	const int comp = a.outComp[o];
	const size_t cb = (size_t)cell * 4 * (N * N);
#pragma unroll
	for (int s = 0; s < E; s++) {
		const int ln = unpad<N>(p + G * s);
		if (ln < 0) continue;
		const int e = ln * N + kn;
		double vr = a.outScale * (double)v[s].x, vi = a.outScale * (double)v[s].y;
		if (a.Wcb && comp > 0) {
			// magnetic rotation 2 (B x w) (src/simulation.c:74-98)
			const double* B = a.Bext + 3 * cell;
			const int c1 = comp == 1 ? 2 : (comp == 2 ? 3 : 1);
			const int c2 = comp == 1 ? 3 : (comp == 2 ? 1 : 2);
			const int b1 = comp == 1 ? 1 : (comp == 2 ? 2 : 0);
			const int b2 = comp == 1 ? 2 : (comp == 2 ? 0 : 1);
			// comp 1: 2(B1 w3 - B2 w2); comp 2: 2(B2 w1 - B0 w3); comp 3: 2(B0 w2 - B1 w1)
			const double2 wa = a.Wcb[cb + (size_t)c2 * (N * N) + e];
			const double2 wb = a.Wcb[cb + (size_t)c1 * (N * N) + e];
			vr += 2 * (B[b1] * wa.x - B[b2] * wb.x);
			vi += 2 * (B[b1] * wa.y - B[b2] * wb.y);
		}
		double2 r = make_double2(a.cf * vr, a.cf * vi);
		if (a.base) {
			const double2 b = a.base[cb + (size_t)comp * (N * N) + e];
			r = make_double2(b.x + r.x, b.y + r.y);
		}
		a.out[cb + (size_t)comp * (N * N) + e] = r;
	}
}

// ---------------------------------------------------------------------------------------------
// Partial sums over term ranges, buf[o][c][e] (c < nch): buf[o][0][e] = sum_c buf[o][c][e], added in
// chunk order (the order k_combine adds them in when it is given the chunks)
template <typename R>
__global__ void k_sum_chunks(cx<R>* __restrict__ buf, int nch, long long nouter, long long PPe)
{
	const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nouter * PPe) return;
	const long long o = i / PPe, e = i % PPe;
	cx<R>* p = buf + o * nch * PPe + e;
	cx<R> s = p[0];
	for (int c = 1; c < nch; c++) s = cadd<R>(s, p[(long long)c * PPe]);
	p[0] = s;
}

// ---------------------------------------------------------------------------------------------
// out = s (a - b) over n complex values: the collision operator of a midpoint step recovered from
// its midpoint state a = b + dt/2 C(b) and start state b, with s = 2 / dt
__global__ void k_scaled_diff(const double2* __restrict__ a, const double2* __restrict__ b, double s,
	double2* __restrict__ out, long long n)
{
	const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const double2 x = a[i], y = b[i];
	out[i] = make_double2(s * (x.x - y.x), s * (x.y - y.y));
}

// ---------------------------------------------------------------------------------------------
// N x N transforms of the state (src/simulation.c:224-283): physical real -> Fourier / N^2
// (full complex FFT of the real data), and Fourier -> unnormalised backward FFT -> real part.
// XG: exchange through the per-workgroup slot xs + blockIdx.x * Xpose::SIZE (N = 128)
template <int N, int T, bool XG>
__global__ __launch_bounds__(T) void k_phys2fourier(const double* __restrict__ in, double2* __restrict__ out,
	const double2* tw256, int ngrid, double2* xs)
{
	using GE = Geo<N, T>;
	constexpr int E = GE::E, G = GE::G;
	__shared__ cx<double> lds[XG ? 1 : Xpose<double, GE>::SIZE];
	__shared__ cx<double> tw[TwTab<N, T>::SIZE];
	load_twiddles<double, N, T>(tw, tw256);
	const int gi = blockIdx.x;
	if (gi >= ngrid) return;
	const int line = threadIdx.x / G, p = threadIdx.x % G;
	const double* src = in + (size_t)gi * (N * N);
	Xpose<double, GE> X{XG ? xs + (size_t)blockIdx.x * Xpose<double, GE>::SIZE : lds};
	cx<double> v[E];
#pragma unroll
	for (int s = 0; s < E; s++) v[s] = make_double2(src[(p + G * s) * N + line], 0.0);
	__syncthreads();
	fft2_reg<double, GE>(v, X, tw, line, p);
	const double sc = 1.0 / ((double)N * N);
	double2* dst = out + (size_t)gi * (N * N);
	const int k0 = E * brev<G>(p);
#pragma unroll
	for (int u = 0; u < E; u++) dst[line * N + k0 + u] = make_double2(v[u].x * sc, v[u].y * sc);
}

template <int N, int T, bool XG>
__global__ __launch_bounds__(T) void k_fourier2phys(const double2* __restrict__ in, double* __restrict__ out,
	const double2* tw256, int ngrid, double2* xs)
{
	using GE = Geo<N, T>;
	constexpr int E = GE::E, G = GE::G;
	__shared__ cx<double> lds[XG ? 1 : Xpose<double, GE>::SIZE];
	__shared__ cx<double> tw[TwTab<N, T>::SIZE];
	load_twiddles<double, N, T>(tw, tw256);
	const int gi = blockIdx.x;
	if (gi >= ngrid) return;
	const int line = threadIdx.x / G, p = threadIdx.x % G;
	const double2* src = in + (size_t)gi * (N * N);
	Xpose<double, GE> X{XG ? xs + (size_t)blockIdx.x * Xpose<double, GE>::SIZE : lds};
	cx<double> v[E];
	const int k0 = E * brev<G>(p);
#pragma unroll
	for (int u = 0; u < E; u++) v[u] = src[line * N + k0 + u];
	__syncthreads();
	ifft2_reg<double, GE>(v, X, tw, line, p);
	double* dst = out + (size_t)gi * (N * N);
#pragma unroll
	for (int s = 0; s < E; s++) dst[(p + G * s) * N + line] = v[s].x;
}

// ---------------------------------------------------------------------------------------------
// Slope-limited upwind transport along x with periodic boundaries (src/finite_volume.c:35-104),
// applied to every (component, l, k) line of a [ncell][4][N][N] state.  Velocity A = vgrid[l].
// One thread per (line element, chunk of cells); the arithmetic follows the reference's
// expression order with contraction disabled, so results are bit-identical to the C code.
__device__ __forceinline__ double minmod_d(double a, double b)
{
	return 0.5 * (copysign(1.0, a) + copysign(1.0, b)) * fmin(fabs(a), fabs(b));
}

__global__ __launch_bounds__(256) void k_transport_periodic(const double* __restrict__ Un, double* __restrict__ Un1,
	int N, int ncell, int chunk, double h, double dt, double L)
{
#pragma clang fp contract(off)
	const int NN = N * N;
	const long long nel = 4LL * NN;
	const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	const int nchunks = (ncell + chunk - 1) / chunk;
	if (gid >= nel * nchunks) return;
	const int e = (int)(gid % nel);          // component * N*N + l*N + k
	const int ch = (int)(gid / nel);
	const int l = (e % NN) / N;
	const double scl = (2 * L) / N;
	const double A = l < N / 2 ? scl * l : scl * (l - N);
	const double Aabs = fabs(A);
	const double nu = dt * A / h;
	const double S = 0.5 * A * (copysign(1.0, nu) - nu);
	const double Sabs = 0.5 * Aabs * (copysign(1.0, nu) - nu);
	const double scale = dt / (2 * h);
	const int c0 = ch * chunk;
	const int c1 = min(ncell, c0 + chunk);
	auto U = [&](int c) -> double {
		c %= ncell;
		if (c < 0) c += ncell;
		return Un[(size_t)c * nel + e];
	};
	// sliding window U[c-2 .. c+2]
	double um2 = U(c0 - 2), um1 = U(c0 - 1), u0 = U(c0), up1 = U(c0 + 1), up2 = U(c0 + 2);
	for (int c = c0; c < c1; c++) {
		const double am2 = um1 - um2, am1 = u0 - um1, a0 = up1 - u0, ap1 = up2 - up1;
		const double bm1 = minmod_d(am1, am2), b0 = minmod_d(a0, am1), bp1 = minmod_d(ap1, a0);
		const double r = u0 + scale * (
			-A * (up1 - um1) + Aabs * (up1 - 2 * u0 + um1)
			- S * (bp1 - bm1) + Sabs * (bp1 - 2 * b0 + bm1));
		Un1[(size_t)c * nel + e] = r;
		um2 = um1; um1 = u0; u0 = up1; up1 = up2;
		up2 = U(c + 3);
	}
}

// Slope-limited transport with Dirichlet ends (src/finite_volume.c:110-170) over all lines
// (component, l, k) of ncell physical states, one thread per line and chunk of cells, the same
// expression order as the single-line host restatement (contraction off, so bit-identical):
// h beta_j = minmod(alpha_j, alpha_{j-1}) inside, 0 at both ends; the upwind end cell keeps its
// value, the downwind end cell takes the one-sided update.
__global__ __launch_bounds__(256) void k_transport_dirichlet(const double* __restrict__ Un, double* __restrict__ Un1,
	int N, int ncell, int chunk, double h, double dt, double L)
{
#pragma clang fp contract(off)
	const int NN = N * N;
	const long long nel = 4LL * NN;
	const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	const int nchunks = (ncell + chunk - 1) / chunk;
	if (gid >= nel * nchunks) return;
	const int e = (int)(gid % nel);
	const int ch = (int)(gid / nel);
	const int l = (e % NN) / N;
	const double scl = (2 * L) / N;
	const double A = l < N / 2 ? scl * l : scl * (l - N);
	const double Aabs = fabs(A);
	const double nu = dt * A / h;
	const double S = 0.5 * A * (copysign(1.0, nu) - nu);
	const double Sabs = 0.5 * Aabs * (copysign(1.0, nu) - nu);
	const double scale = dt / (2 * h);
	const int n = ncell;
	auto U = [&](int c) -> double { return Un[(size_t)c * nel + e]; };
	auto HB = [&](int j) -> double { return (j <= 0 || j >= n - 1) ? 0.0 : minmod_d(U(j + 1) - U(j), U(j) - U(j - 1)); };
	const int c0 = ch * chunk, c1 = min(n, c0 + chunk);
	for (int c = c0; c < c1; c++) {
		double r;
		if (c == 0) {
			r = A >= 0 ? U(0) : U(0) + 2 * scale * (-A * (U(1) - U(0)) - S * (HB(1) - HB(0)));
		} else if (c == n - 1) {
			r = A >= 0 ? U(n - 1) + 2 * scale * (-A * (U(n - 1) - U(n - 2)) - S * (HB(n - 1) - HB(n - 2))) : U(n - 1);
		} else {
			const double bm = HB(c - 1), b0 = HB(c), bp = HB(c + 1);
			r = U(c) + scale * (
				-A * (U(c + 1) - U(c - 1)) + Aabs * (U(c + 1) - 2 * U(c) + U(c - 1))
				- S * (bp - bm) + Sabs * (bp - 2 * b0 + bm));
		}
		Un1[(size_t)c * nel + e] = r;
	}
}

// Trace fluxes of the Maxwell step on the Maxwell velocity grid (src/finite_volume.c:186-214,
// src/simulation.c:905-918, 1053-1059): s[0] = leftward flux of the trace of cell 0, s[1] =
// rightward flux of the trace of cell ncell-1, s[2] = 1 / rightward flux of WL's trace, s[3] =
// 1 / leftward flux of WR's trace.  The Maxwell grid drops the zero x-velocity row and holds the
// -L row twice: Maxwell index i < NN/2 is standard index i + N with velocity (l+1)*2L/N, i >= NN/2
// is standard index i with velocity (l-N)*2L/N (l = i / N).  The products are formed by all
// threads; each sum then runs in index order in one lane, so the rounding is the reference's.
__global__ __launch_bounds__(256) void k_maxwell_flux(const double* __restrict__ W, int N, int ncell, double L,
	const double* __restrict__ WL, const double* __restrict__ WR, double* __restrict__ s)
{
#pragma clang fp contract(off)
	constexpr int TILE = 1024;
	__shared__ double prod[4][TILE];
	const int NN = N * N, half = NN / 2;
	const double scl = (2 * L) / N;
	const double* Wlast = W + (size_t)(ncell - 1) * 4 * NN;
	double acc = 0.0;
	for (int base = 0; base < half; base += TILE) {
		for (int t = threadIdx.x; t < 4 * TILE; t += blockDim.x) {
			const int w = t / TILE, i = base + t % TILE;
			if (i >= half) continue;
			const bool right = w == 1 || w == 2;                  // rightward sums run over Maxwell i < NN/2
			const int l = (right ? i : half + i) / N;
			const double A = l < N / 2 ? scl * (l + 1) : scl * (l - N);
			const double* tr = w == 0 ? W : w == 1 ? Wlast : w == 2 ? WL : WR;
			const double u = tr[right ? i + N : half + i];
			prod[w][t % TILE] = right ? A * u : fabs(A) * u;
		}
		__syncthreads();
		if (threadIdx.x < 4) {
			const int cnt = min(TILE, half - base);
			for (int i = 0; i < cnt; i++) acc += prod[threadIdx.x][i];
		}
		__syncthreads();
	}
	if (threadIdx.x < 4) s[threadIdx.x] = threadIdx.x >= 2 ? 1.0 / acc : acc;
}

// Half transport of ncell cells with Maxwell reflection at both ends (src/finite_volume.c:241-323
// on the Maxwell grid, src/simulation.c:1020-1029).  Thread = (component, velocity), one chunk of
// cells.  Standard row r >= 1 is Maxwell row r-1 (r < N/2) or r (r >= N/2) with velocity
// vgrid[r]; the zero x-velocity row r = 0 is not transported.  The ghost cells V(-2) = V(-1) and
// V(n) = V(n+1) repeat the end cell on the outgoing side; on the incoming side they mix the
// reflected end cell (Maxwell index NN-1-i, standard row N-r, column N-1-k) with the normalised
// boundary state: (1 - lambda) U_refl + (lambda flux) (W_b * norm).  The Maxwell +L copy of the
// -L row is only read (as a reflection), never written, as in the reference.
__global__ __launch_bounds__(256) void k_transport_maxwell(const double* __restrict__ Un, double* __restrict__ Un1,
	int N, int ncell, int chunk, double h, double dt, double L, double lambda, const double* __restrict__ WL,
	const double* __restrict__ WR, const double* __restrict__ s)
{
#pragma clang fp contract(off)
	const int NN = N * N;
	const long long nel = 4LL * NN;
	const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	const int nchunks = (ncell + chunk - 1) / chunk;
	if (gid >= nel * nchunks) return;
	const int e = (int)(gid % nel);
	const int ch = (int)(gid / nel);
	const int j = e / NN, q = e % NN, r = q / N, k = q % N;
	const int n = ncell;
	const int c0 = ch * chunk, c1 = min(n, c0 + chunk);
	if (r == 0) {
		for (int c = c0; c < c1; c++) Un1[(size_t)c * nel + e] = Un[(size_t)c * nel + e];
		return;
	}
	const double scl = (2 * L) / N;
	const double A = r < N / 2 ? scl * r : scl * (r - N);
	const double Aabs = fabs(A);
	const double nu = dt * A / h;
	const double S = 0.5 * A * (copysign(1.0, nu) - nu);
	const double Sabs = 0.5 * Aabs * (copysign(1.0, nu) - nu);
	const double scale = dt / (2 * h);
	const int qr = j * NN + (N - r) * N + (N - 1 - k);
	double gm, gp;
	if (r < N / 2) {
		const double ub = WL[e] * s[2];
		gm = (1 - lambda) * Un[qr] + lambda * s[0] * ub;
		gp = Un[(size_t)(n - 1) * nel + e];
	} else {
		const double ub = WR[e] * s[3];
		gm = Un[e];
		gp = (1 - lambda) * Un[(size_t)(n - 1) * nel + qr] + lambda * s[1] * ub;
	}
	auto V = [&](int c) -> double { return c < 0 ? gm : c >= n ? gp : Un[(size_t)c * nel + e]; };
	auto HB = [&](int c) -> double { return minmod_d(V(c + 1) - V(c), V(c) - V(c - 1)); };
	for (int c = c0; c < c1; c++) {
		const double vm = V(c - 1), v0 = V(c), vp = V(c + 1);
		const double bm = HB(c - 1), b0 = HB(c), bp = HB(c + 1);
		Un1[(size_t)c * nel + e] = v0 + scale * (
			-A * (vp - vm) + Aabs * (vp - 2 * v0 + vm)
			- S * (bp - bm) + Sabs * (bp - 2 * b0 + bm));
	}
}

// One line of n cells with velocity A (the single-line entry point of the reference,
// src/finite_volume.c:52-104); same arithmetic as k_transport_periodic.
__global__ __launch_bounds__(256) void k_transport_line(const double* __restrict__ Un, double* __restrict__ Un1,
	int n, double A, double h, double dt)
{
#pragma clang fp contract(off)
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= n) return;
	const double Aabs = fabs(A);
	const double nu = dt * A / h;
	const double S = 0.5 * A * (copysign(1.0, nu) - nu);
	const double Sabs = 0.5 * Aabs * (copysign(1.0, nu) - nu);
	const double scale = dt / (2 * h);
	auto U = [&](int i) -> double { i %= n; if (i < 0) i += n; return Un[i]; };
	const double um2 = U(c - 2), um1 = U(c - 1), u0 = U(c), up1 = U(c + 1), up2 = U(c + 2);
	const double am2 = um1 - um2, am1 = u0 - um1, a0 = up1 - u0, ap1 = up2 - up1;
	const double bm1 = minmod_d(am1, am2), b0 = minmod_d(a0, am1), bp1 = minmod_d(ap1, a0);
	Un1[c] = u0 + scale * (
		-A * (up1 - um1) + Aabs * (up1 - 2 * u0 + um1)
		- S * (bp1 - bm1) + Sabs * (bp1 - 2 * b0 + bm1));
}

// ---------------------------------------------------------------------------------------------
// FP64 conservation correction of a collision step, one workgroup per cell.
// dW = Wnew - Wold (physical).  The trace part is corrected by
//   dW0 -= |w0| (lambda . psi),  psi = (1, v_x, v_y, |v|^2/2),
// with lambda solving G lambda = m, G_jk = sum |w0| psi_j psi_k, m_j = sum psi_j dW0 (weighted
// least squares: the smallest |w0|^-1-weighted change that zeroes the selected moments of dW0);
// each spin component by dW_c -= (sum dW_c / sum |w0|) |w0|.  mask bit 0: mass and spin,
// bit 1: momentum, bit 2: energy (moments as in visualization/wignerMoments.m:15-37).
__device__ __forceinline__ double block_sum_256(double x, double* sh)
{
	for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
	const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
	__syncthreads();
	if (l == 0) sh[w] = x;
	__syncthreads();
	double r = 0;
	for (int i = 0; i < (int)(blockDim.x >> 6); i++) r += sh[i];
	return r;
}

__global__ __launch_bounds__(256) void k_moment_correct(const double* __restrict__ Wold, double* __restrict__ Wnew,
	int N, double L, int mask)
{
	__shared__ double sh[4];
	__shared__ double sol[8];
	const int NN = N * N;
	const size_t cb = (size_t)blockIdx.x * 4 * NN;
	const double sc = (2 * L) / N;
	// accumulate the Gram matrix (10 entries, [0] = sum |w0|), the trace moments (4), spin sums (3)
	double acc[17];
	for (int i = 0; i < 17; i++) acc[i] = 0;
	for (int e = threadIdx.x; e < NN; e += blockDim.x) {
		const int l = e / N, k = e % N;
		const double vx = l < N / 2 ? sc * l : sc * (l - N);
		const double vy = k < N / 2 ? sc * k : sc * (k - N);
		const double ps[4] = {1.0, vx, vy, 0.5 * (vx * vx + vy * vy)};
		const double w = fabs(Wold[cb + e]);
		const double d0 = Wnew[cb + e] - Wold[cb + e];
		int t = 0;
		for (int a = 0; a < 4; a++)
			for (int b = a; b < 4; b++) acc[t++] += w * ps[a] * ps[b];
		for (int a = 0; a < 4; a++) acc[10 + a] += ps[a] * d0;
		for (int c = 1; c < 4; c++) acc[13 + c] += Wnew[cb + (size_t)c * NN + e] - Wold[cb + (size_t)c * NN + e];
	}
	double tot[17];
	for (int i = 0; i < 17; i++) tot[i] = block_sum_256(acc[i], sh);
	if (threadIdx.x == 0) {
		// selected constraints
		int idx[4], n = 0;
		if (mask & 1) idx[n++] = 0;
		if (mask & 2) { idx[n++] = 1; idx[n++] = 2; }
		if (mask & 4) idx[n++] = 3;
		double G[4][4], m[4], lam[4] = {0, 0, 0, 0};
		auto gram = [&](int a, int b) {
			if (a > b) { int t = a; a = b; b = t; }
			const int off[4] = {0, 4, 7, 9};
			return tot[off[a] + (b - a)];
		};
		for (int i = 0; i < n; i++) {
			for (int j = 0; j < n; j++) G[i][j] = gram(idx[i], idx[j]);
			m[i] = tot[10 + idx[i]];
		}
		// Gaussian elimination with partial pivoting on the n x n system
		for (int c = 0; c < n; c++) {
			int pv = c;
			for (int r = c + 1; r < n; r++) if (fabs(G[r][c]) > fabs(G[pv][c])) pv = r;
			for (int j = 0; j < n; j++) { double t = G[c][j]; G[c][j] = G[pv][j]; G[pv][j] = t; }
			{ double t = m[c]; m[c] = m[pv]; m[pv] = t; }
			for (int r = c + 1; r < n; r++) {
				const double f = G[c][c] != 0 ? G[r][c] / G[c][c] : 0;
				for (int j = c; j < n; j++) G[r][j] -= f * G[c][j];
				m[r] -= f * m[c];
			}
		}
		for (int c = n - 1; c >= 0; c--) {
			double t = m[c];
			for (int j = c + 1; j < n; j++) t -= G[c][j] * lam[idx[j]];
			lam[idx[c]] = G[c][c] != 0 ? t / G[c][c] : 0;
		}
		for (int i = 0; i < 4; i++) sol[i] = lam[i];
		const double wsum = tot[0];
		for (int c = 1; c < 4; c++) sol[3 + c] = ((mask & 1) && wsum != 0) ? tot[13 + c] / wsum : 0;
	}
	__syncthreads();
	for (int e = threadIdx.x; e < NN; e += blockDim.x) {
		const int l = e / N, k = e % N;
		const double vx = l < N / 2 ? sc * l : sc * (l - N);
		const double vy = k < N / 2 ? sc * k : sc * (k - N);
		const double w = fabs(Wold[cb + e]);
		Wnew[cb + e] -= w * (sol[0] + sol[1] * vx + sol[2] * vy + sol[3] * 0.5 * (vx * vx + vy * vy));
		for (int c = 1; c < 4; c++) Wnew[cb + (size_t)c * NN + e] -= sol[3 + c] * w;
	}
}

// Per-cell diagnostics of physical states [ncell][4][N][N] (one workgroup of 256 per cell):
//   out[0..3]  sum_v w_c dv                    (mass and spin densities)
//   out[4..5]  sum_v v_x w0 dv, sum_v v_y w0 dv  (momentum)
//   out[6]     sum_v |v|^2/2 w0 dv             (energy)
//   out[7]     sum_v [s(l+) + s(l-)] dv,  s(p) = -(p ln p + (1-p) ln(1-p)) on (0, 1), else 0
//   out[8], out[9]  min and max eigenvalue over v
// with the eigenvalues l+- = w0 +- |(w1, w2, w3)| of the 2x2 matrix w0 I + w . sigma, dv = (2L/N)^2
// (visualization/wignerMoments.m:15-37, entropy.m, visualizeHomBoltzmann.m:26-45, 98-105).
__device__ __forceinline__ double bin_entropy(double p)
{
	return (p > 0.0 && p < 1.0) ? -(p * log(p) + (1.0 - p) * log1p(-p)) : 0.0;
}

__global__ __launch_bounds__(256) void k_cell_diag(const double* __restrict__ W, int N, double L,
	double* __restrict__ out)
{
	__shared__ double sh[4];
	const int NN = N * N;
	const double* w = W + (size_t)blockIdx.x * 4 * NN;
	const double sc = (2 * L) / N;
	double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
	double lmin = INFINITY, lmax = -INFINITY;
	for (int e = threadIdx.x; e < NN; e += blockDim.x) {
		const int l = e / N, k = e % N;
		const double vx = l < N / 2 ? sc * l : sc * (l - N);
		const double vy = k < N / 2 ? sc * k : sc * (k - N);
		const double w0 = w[e], w1 = w[NN + e], w2 = w[2 * NN + e], w3 = w[3 * NN + e];
		acc[0] += w0;
		acc[1] += w1;
		acc[2] += w2;
		acc[3] += w3;
		acc[4] += vx * w0;
		acc[5] += vy * w0;
		acc[6] += 0.5 * (vx * vx + vy * vy) * w0;
		const double r = sqrt(w1 * w1 + w2 * w2 + w3 * w3);
		acc[7] += bin_entropy(w0 + r) + bin_entropy(w0 - r);
		lmin = fmin(lmin, w0 - r);
		lmax = fmax(lmax, w0 + r);
	}
	const double dv = sc * sc;
	for (int i = 0; i < 8; i++) {
		const double t = block_sum_256(acc[i], sh);
		if (threadIdx.x == 0) out[(size_t)blockIdx.x * 10 + i] = t * dv;
	}
	for (int o = 32; o > 0; o >>= 1) {
		lmin = fmin(lmin, __shfl_xor(lmin, o));
		lmax = fmax(lmax, __shfl_xor(lmax, o));
	}
	__syncthreads();
	if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = lmin;
	__syncthreads();
	if (threadIdx.x == 0) out[(size_t)blockIdx.x * 10 + 8] = fmin(fmin(sh[0], sh[1]), fmin(sh[2], sh[3]));
	__syncthreads();
	if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = lmax;
	__syncthreads();
	if (threadIdx.x == 0) out[(size_t)blockIdx.x * 10 + 9] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

// ---------------------------------------------------------------------------------------------
// Packed component pairs for the paired path: Wp[cell] = {W0 + i W1, W2 + i W3, W1} (n = ncell * 3 * NN
// elements).  The padded transform of a packed slot (and of psi .* it) is the pair of transforms of
// its two components in its real and imaginary parts when both are real.
__global__ __launch_bounds__(256) void k_pack_pairs(const double2* __restrict__ W, double2* __restrict__ Wp, int NN,
	long long n)
{
	const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const long long cell = i / (3 * NN);
	const int r = (int)(i % (3 * NN));
	const int pr = r / NN, e = r % NN;
	if (pr == 2) {
		Wp[i] = W[(cell * 4 + 1) * NN + e];
		return;
	}
	const double2 x = W[(cell * 4 + 2 * pr) * NN + e], y = W[(cell * 4 + 2 * pr + 1) * NN + e];
	Wp[i] = make_double2(x.x - y.y, x.y + y.x);
}

// Unpacking of packed I1-type results: in [ncell][npk][nch][PP] -> out [ncell][nout][nch][PP],
// the real part of packed slot k into slot dst[k][0], the imaginary part into dst[k][1] (each with
// imaginary part zero; -1: dropped).  n = ncell * npk * nch * PP.
template <typename R>
__global__ __launch_bounds__(256) void k_unpack_real(UnpackArgs a, long long n, long long PP)
{
	const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const long long e = i % PP, t = i / PP;
	const int ch = (int)(t % a.nch), k = (int)((t / a.nch) % a.npk);
	const long long cell = t / ((long long)a.nch * a.npk);
	const cx<R> z = ((const cx<R>*)a.in)[i];
	for (int c = 0; c < 2; c++) {
		const int d = a.dst[k][c];
		if (d >= 0)
			((cx<R>*)a.out)[((cell * a.nout + d) * a.nch + ch) * PP + e] = mk<R>(c ? z.y : z.x, R(0));
	}
}

// Largest |real part| and |imaginary part| of n complex values: out[block] = (max |re|, max |im|)
// over the block's grid-stride share.
template <typename R>
__global__ __launch_bounds__(256) void k_re_im_max(const cx<R>* __restrict__ g, long long n, double* __restrict__ out)
{
	__shared__ double sh[2][4];
	double mr = 0, mi = 0;
	for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
		const cx<R> z = g[i];
		mr = fmax(mr, fabs((double)z.x));
		mi = fmax(mi, fabs((double)z.y));
	}
	for (int o = 32; o > 0; o >>= 1) {
		mr = fmax(mr, __shfl_xor(mr, o));
		mi = fmax(mi, __shfl_xor(mi, o));
	}
	if ((threadIdx.x & 63) == 0) {
		sh[0][threadIdx.x >> 6] = mr;
		sh[1][threadIdx.x >> 6] = mi;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		out[2 * blockIdx.x] = fmax(fmax(sh[0][0], sh[0][1]), fmax(sh[0][2], sh[0][3]));
		out[2 * blockIdx.x + 1] = fmax(fmax(sh[1][0], sh[1][1]), fmax(sh[1][2], sh[1][3]));
	}
}

// ---------------------------------------------------------------------------------------------
// Fermi-Dirac equilibrium diagnostics (visualization/asymptoticFD.m, calculateFDparams.m,
// createFD.m, relentropy.m, visualizeHomBoltzmann.m:114-149, visualizeInhomBoltzmann.m:163-205).

// Dilogarithm Li2(x) for x <= 0.  For -1 <= x <= 0 the Bernoulli series in u = -ln(1 - x),
// Li2(x) = sum_k B_k u^(k+1) / (k+1)!  (|u| <= ln 2, terms up to k = 20 reach double precision);
// for x < -1 the inversion Li2(x) = -pi^2/6 - ln^2(-x)/2 - Li2(1/x).  ('t Hooft and Veltman,
// Nucl. Phys. B153 (1979) 365.)
__device__ double li2_nonpos(double x)
{
	double add = 0.0, sgn = 1.0;
	if (x < -1.0) {
		const double l = log(-x);
		add = -M_PI * M_PI / 6.0 - 0.5 * l * l;
		sgn = -1.0;
		x = 1.0 / x;
	}
	const double u = -log1p(-x), u2 = u * u;
	// B_k / (k+1)! for k = 2, 4, ..., 20
	constexpr double c[10] = {1.0 / 36.0, -1.0 / 3600.0, 1.0 / 211680.0, -1.0 / 10886400.0, 1.0 / 526901760.0,
		-4.0647616451442256e-11, 8.921691020456452e-13, -1.9939295860721074e-14, 4.518980029619918e-16,
		-1.0356517612181247e-17};
	double t = c[9];
	for (int i = 8; i >= 0; i--) t = c[i] + u2 * t;
	return add + sgn * (u - 0.25 * u2 + u * u2 * t);
}

// Energy of the two-species 2D Fermi-Dirac gas with densities l1, l2 at inverse temperature beta
// (calculateFDparams.m:65-83): the density equations 2 pi / beta ln(1 + z) = l fix the fugacities,
// z = expm1(beta l / 2 pi), and E = -2 pi / beta^2 (Li2(-z1) + Li2(-z2)).
__device__ double fd_energy(double beta, double l1, double l2)
{
	const double k = beta / (2.0 * M_PI);
	return -(2.0 * M_PI) / (beta * beta) * (li2_nonpos(-expm1(l1 * k)) + li2_nonpos(-expm1(l2 * k)));
}

// beta with fd_energy(beta) = E.  E decreases monotonically from +inf (beta -> 0) to the ground
// state (l1^2 + l2^2) / 4 pi (beta -> inf), so a bracket in beta and bisection in log(beta) find
// the unique root; returns 0 (status 1) when there is none.
__device__ double fd_solve_beta(double l1, double l2, double E, int* status)
{
	*status = 1;
	if (!(l1 > 0.0) || !(l2 >= l1) || !(E > (l1 * l1 + l2 * l2) / (4.0 * M_PI)) || !(E < INFINITY)) return 0.0;
	const double b0 = (l1 + l2) / E;   // classical limit, exact as beta -> 0
	double lo = b0, hi = b0;
	for (int i = 0; i < 1100 && fd_energy(lo, l1, l2) < E; i++) lo *= 0.5;
	for (int i = 0; i < 1100 && fd_energy(hi, l1, l2) > E; i++) hi *= 2.0;
	if (!(fd_energy(lo, l1, l2) >= E) || !(fd_energy(hi, l1, l2) <= E)) return 0.0;
	for (int it = 0; it < 200 && hi > lo * (1.0 + 4e-16); it++) {
		const double mid = sqrt(lo * hi);
		if (fd_energy(mid, l1, l2) > E) lo = mid;
		else hi = mid;
	}
	*status = 0;
	return sqrt(lo * hi);
}

// Per cell: the Fermi-Dirac state with the cell's conserved moments (asymptoticFD.m).  Moments as
// wignerMoments.m:15-37 (spin density matrix rho = w0 + w.sigma, velocity, intrinsic energy about
// it); eigenvalues l1 <= l2 of rho with eigenvectors -n, +n; then beta, mu_i = ln(z_i) / beta and
// W_FD(v) = f1 P- + f2 P+, f_i = 1 / (exp(beta (|v - vel|^2 / 2 - mu_i)) + 1), P+- = (1 +- n.sigma) / 2,
// i.e. Pauli components ((f1 + f2) / 2, (f2 - f1) / 2 n).
// params[cell][12] = beta, mu1, mu2, vel_x, vel_y, n_x, n_y, n_z, l1, l2, en, status (0: fitted,
// 1: no Fermi-Dirac state has these moments, W_FD = 0).
__global__ __launch_bounds__(256) void k_fermi_dirac(const double* __restrict__ W, int N, double L,
	double* __restrict__ params, double* __restrict__ WFD)
{
	__shared__ double sh[4];
	__shared__ double prm[12];
	const int NN = N * N;
	const double* w = W + (size_t)blockIdx.x * 4 * NN;
	const double sc = (2 * L) / N, dv = sc * sc;
	double acc[6] = {0, 0, 0, 0, 0, 0};
	for (int e = threadIdx.x; e < NN; e += blockDim.x) {
		const int l = e / N, k = e % N;
		const double vx = l < N / 2 ? sc * l : sc * (l - N);
		const double vy = k < N / 2 ? sc * k : sc * (k - N);
		const double w0 = w[e];
		acc[0] += w0;
		acc[1] += w[NN + e];
		acc[2] += w[2 * NN + e];
		acc[3] += w[3 * NN + e];
		acc[4] += vx * w0;
		acc[5] += vy * w0;
	}
	double tot[6];
	for (int i = 0; i < 6; i++) tot[i] = block_sum_256(acc[i], sh);
	const double velx = tot[4] / tot[0], vely = tot[5] / tot[0];
	double ein = 0;
	for (int e = threadIdx.x; e < NN; e += blockDim.x) {
		const int l = e / N, k = e % N;
		const double dx = (l < N / 2 ? sc * l : sc * (l - N)) - velx;
		const double dy = (k < N / 2 ? sc * k : sc * (k - N)) - vely;
		ein += 0.5 * (dx * dx + dy * dy) * w[e];
	}
	ein = block_sum_256(ein, sh);
	if (threadIdx.x == 0) {
		const double m0 = tot[0] * dv, mx = tot[1] * dv, my = tot[2] * dv, mz = tot[3] * dv;
		const double r = sqrt(mx * mx + my * my + mz * mz);
		const double l1 = m0 - r, l2 = m0 + r;
		const double en = ein / tot[0];                 // normalised intrinsic energy
		int st;
		const double beta = fd_solve_beta(l1, l2, en * (l1 + l2), &st);
		const double k = beta / (2.0 * M_PI);
		prm[0] = beta;
		prm[1] = st ? 0.0 : log(expm1(l1 * k)) / beta;
		prm[2] = st ? 0.0 : log(expm1(l2 * k)) / beta;
		prm[3] = velx;
		prm[4] = vely;
		prm[5] = r > 0 ? mx / r : 0.0;
		prm[6] = r > 0 ? my / r : 0.0;
		prm[7] = r > 0 ? mz / r : 0.0;
		prm[8] = l1;
		prm[9] = l2;
		prm[10] = en;
		prm[11] = st;
	}
	__syncthreads();
	if (threadIdx.x < 12) params[(size_t)blockIdx.x * 12 + threadIdx.x] = prm[threadIdx.x];
	if (!WFD) return;
	double* o = WFD + (size_t)blockIdx.x * 4 * NN;
	const bool ok = prm[11] == 0.0;
	for (int e = threadIdx.x; e < NN; e += blockDim.x) {
		const int l = e / N, k = e % N;
		const double dx = (l < N / 2 ? sc * l : sc * (l - N)) - prm[3];
		const double dy = (k < N / 2 ? sc * k : sc * (k - N)) - prm[4];
		const double ek = 0.5 * (dx * dx + dy * dy);
		const double f1 = ok ? 1.0 / (exp(prm[0] * (ek - prm[1])) + 1.0) : 0.0;
		const double f2 = ok ? 1.0 / (exp(prm[0] * (ek - prm[2])) + 1.0) : 0.0;
		const double h = 0.5 * (f2 - f1);
		o[e] = 0.5 * (f1 + f2);
		o[NN + e] = h * prm[5];
		o[2 * NN + e] = h * prm[6];
		o[3 * NN + e] = h * prm[7];
	}
}

// Quantum relative entropy S(rho || tau) of two 2x2 density matrices given by Pauli components
// (relentropy.m): eigenvalues clamped to [eps, 1 - eps] (rho is replaced by its clamped form),
// tr rho (ln rho - ln tau) + tr (1 - rho)(ln(1 - rho) - ln(1 - tau)), in closed form: for
// A = a + b.sigma and B = c + d.sigma, tr AB = 2 (a c + b.d).
__device__ double relentropy2(double a, double b1, double b2, double b3, double c, double d1, double d2, double d3)
{
	const double eps = 2.220446049250313e-16;
	auto clampe = [&](double x) { return x <= 0.0 ? eps : x >= 1.0 ? 1.0 - eps : x; };
	const double rb = sqrt(b1 * b1 + b2 * b2 + b3 * b3), rd = sqrt(d1 * d1 + d2 * d2 + d3 * d3);
	const double mp = clampe(a + rb), mm = clampe(a - rb), np_ = clampe(c + rd), nm = clampe(c - rd);
	const double cosbd = (rb > 0 && rd > 0) ? (b1 * d1 + b2 * d2 + b3 * d3) / (rb * rd) : 0.0;
	const double am = 0.5 * (mp + mm), bm = 0.5 * (mp - mm);          // clamped rho = am + bm b^.sigma
	const double lt0 = 0.5 * (log(np_) + log(nm)), lt1 = 0.5 * (log(np_) - log(nm));
	const double lu0 = 0.5 * (log1p(-np_) + log1p(-nm)), lu1 = 0.5 * (log1p(-np_) - log1p(-nm));
	const double rr = mp * log(mp) + mm * log(mm) + (1 - mp) * log1p(-mp) + (1 - mm) * log1p(-mm);
	const double rt = 2.0 * (am * lt0 + bm * lt1 * cosbd);             // tr rho ln tau
	const double ru = 2.0 * ((1 - am) * lu0 - bm * lu1 * cosbd);       // tr (1 - rho) ln(1 - tau)
	return rr - rt - ru;
}

// Per cell: distance and relative entropy to a Fermi-Dirac state (visualizeHomBoltzmann.m:120-147):
// out[cell][0] = sum_j ||W_j - WFD_j||_1 dv with the matrix 1-norm of the N x N grid (maximum over
// the slow index l of sum_k |.|, as MATLAB's norm(A, 1) on the [k, l] matrix), out[cell][1] =
// sum_v S(W(v) || WFD(v)) dv over the velocities with WFD_0(v) > 1e-6.  fd_stride 0: one state WFD
// for every cell (the global equilibrium); 1: one per cell.
__global__ __launch_bounds__(256) void k_fd_distance(const double* __restrict__ W, const double* __restrict__ WFD,
	int fd_stride, int N, double L, double* __restrict__ out)
{
	__shared__ double sh[4];
	const int NN = N * N;
	const double* w = W + (size_t)blockIdx.x * 4 * NN;
	const double* f = WFD + (size_t)blockIdx.x * fd_stride * 4 * NN;
	const double sc = (2 * L) / N, dv = sc * sc;
	// one wave per component, a lane per column l (and l + 64, ... for N > 64)
	const int j = threadIdx.x >> 6, lane = threadIdx.x & 63;
	double col = 0;
	for (int l = lane; l < N; l += 64) {
		double c = 0;
		for (int k = 0; k < N; k++) c += fabs(w[j * NN + k + N * l] - f[j * NN + k + N * l]);
		col = fmax(col, c);
	}
	for (int o = 32; o > 0; o >>= 1) col = fmax(col, __shfl_xor(col, o));
	double dist = (lane == 0) ? col : 0.0;
	dist = block_sum_256(dist, sh);
	double kl = 0;
	for (int e = threadIdx.x; e < NN; e += blockDim.x)
		if (f[e] > 1e-6)
			kl += relentropy2(w[e], w[NN + e], w[2 * NN + e], w[3 * NN + e], f[e], f[NN + e], f[2 * NN + e], f[3 * NN + e]);
	kl = block_sum_256(kl, sh);
	if (threadIdx.x == 0) {
		out[(size_t)blockIdx.x * 2] = dist * dv;
		out[(size_t)blockIdx.x * 2 + 1] = kl * dv;
	}
}

// Average over the cells, summed in cell order (visualizeInhomBoltzmann.m:80-89).
__global__ __launch_bounds__(256) void k_cell_average(const double* __restrict__ W, int ncell, long long nel,
	double* __restrict__ out)
{
	const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= nel) return;
	double s = 0.0;
	for (int c = 0; c < ncell; c++) s += W[(size_t)c * nel + e];
	out[e] = s / ncell;
}

// y = a * x + b * y  (elementwise, double)
__global__ void k_axpby(const double* __restrict__ x, double* __restrict__ y, double a, double b, long long n)
{
	const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) y[i] = a * x[i] + b * y[i];
}

} // namespace qbf
