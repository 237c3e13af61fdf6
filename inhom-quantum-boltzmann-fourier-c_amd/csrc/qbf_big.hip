// qbf_big.hip -- collision kernels for N = 128 (zero-padded P = 256 grids).
//
// An FP32 256 x 256 complex grid is 512 KiB and an FP64 one 1 MiB: it fits neither in the
// registers nor in the 160 KiB of LDS of a CU.  So at N = 128 a padded grid lives in global
// memory (the per-workgroup slot a.xs + blockIdx.x * P*P, L2 / MALL resident while the workgroup
// works on it) in natural order, element (l, k) at l * P + k (l the slow v_x index, as the
// reference's grid2_t), and the 1024 threads of a workgroup run its line FFTs in sweeps of 64
// lines of 16 lanes x 16 elements: rows (lines along k, contiguous) or columns (lines along l,
// stride P).  Every 2-D transform is a row pass and a column pass with a workgroup barrier
// between them; pointwise factors are applied inside the passes.
//
// The kernels take the argument blocks of the register-resident kernels of smaller N
// (qbf_kernels.hip) with the same meaning; all their P x P grids (spectra, tables, partial sums)
// are in natural order instead of the register-native order.
#include "qbf_kernels.hip"

namespace qbf {
namespace big {
constexpr int N = 128, P = 256, G = 16, E = 16;
// threads per workgroup: 512 (8 waves, up to 256 VGPRs a lane: the line FFTs of the FP64 grids
// and the instance sums of the I3 kernel fit without spills)
template <typename R> constexpr int TB = 512;
constexpr int I3_NI = 4;                    // scratch grids per I3 workgroup (instances of an output)
constexpr int TW = 2 * P + G;               // twiddle table entries (load_twiddles_g)
constexpr size_t PP = (size_t)P * P;
static_assert(2 * N == P && G * E == P, "geometry");

// blocked element u of lane p <-> index E brev(p) + u; interleaved element s <-> p + G s
__device__ __forceinline__ int blk(int p) { return E * brev<G>(p); }

// row of the padded grid that holds row ln of the N x N grid
__device__ __forceinline__ int padrow(int ln) { return ln < N / 2 ? ln : ln + N; }

// Pass A of the pruned FFT2 of pad(pre .* src): the N non-zero rows, FFT along k, stored in
// natural order at row padrow(ln) ^ rx of dst (rx = 0, or 128 to use the rows that the padding
// leaves empty: 64 .. 191).  preKind 0 none, 1 real N x N, 2 complex N x N (src/grid.c:171-188).
template <typename R, int L>
__device__ __forceinline__ void pad_rows(const double2* __restrict__ src, const void* pre, int preKind, int rx,
	cx<R>* dst, const cx<R>* tw, int lt, int p)
{
	const double2* sp = uni(src, 0);
	const double* p1 = uni((const double*)pre, 0);
	const double2* p2 = uni((const double2*)pre, 0);
#pragma unroll 1
	for (int sw = 0; sw < N / L; sw++) {
		const int ln = sw * L + lt;
		cx<R> v[E];
#pragma unroll
		for (int s = 0; s < E; s++) {
			if (padded_nonzero<E>(s)) {
				const int e = ln * N + p + G * s - (s >= E / 4 ? N : 0);
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
		cx<R>* row = dst + (padrow(ln) ^ rx) * P + blk(p);
#pragma unroll
		for (int u = 0; u < E; u++) row[u] = v[u];
	}
}

// column c of a pass-A grid (rows padrow(.) ^ rx), interleaved, padding rows as zeros
template <typename R>
__device__ __forceinline__ void load_padded_col(cx<R>* v, const cx<R>* xs, int rx, int c, int p)
{
#pragma unroll
	for (int s = 0; s < E; s++) v[s] = padded_nonzero<E>(s) ? xs[((p + G * s) ^ rx) * P + c] : czero<R>();
}

} // namespace big

// ---------------------------------------------------------------------------------------------
// out[o] = FFT2(pad(pre_q .* src)), one grid per workgroup (PadFFTArgs as k_padfft)
template <typename R>
__global__ __launch_bounds__(big::TB<R>) void k_padfft_big(PadFFTArgs a)
{
	using namespace big;
	constexpr int T = TB<R>, L = T / G;         // L: lines per sweep
	__shared__ cx<R> tw[TW];
	load_twiddles_g<R, P, G, T>(tw, a.tw256);
	const int o = blockIdx.x;
	const int lt = threadIdx.x / G, p = threadIdx.x % G;
	cx<R>* xs = (cx<R>*)a.xs + (size_t)blockIdx.x * PP;
	cx<R>* out = (cx<R>*)a.out + (size_t)o * PP;
	const int slot = o % a.nslot;
	const int qq = (o / a.nslot) % a.nq;
	const int cell = o / (a.nslot * a.nq);
	const double2* src = a.src + ((size_t)cell * a.nsrc + a.srcOfSlot[slot]) * (N * N);
	const size_t poff = (size_t)(a.q0 + qq) * (N * N);
	const void* pre = a.preKind == 1 ? (const void*)((const double*)a.pre + poff)
		: a.preKind == 2 ? (const void*)((const double2*)a.pre + poff) : nullptr;
	__syncthreads();
	pad_rows<R, L>(src, pre, a.preKind, 0, xs, tw, lt, p);
	__syncthreads();
#pragma unroll 1
	for (int sw = 0; sw < P / L; sw++) {
		const int c = sw * L + lt;
		cx<R> v[E];
		load_padded_col<R>(v, xs, 0, c, p);
		line_fwd<R, P, G, E>(v, p, tw);
		const int r0 = blk(p);
#pragma unroll
		for (int u = 0; u < E; u++) out[(r0 + u) * P + c] = v[u];
	}
}

// ---------------------------------------------------------------------------------------------
// out[cell][inst][chunk] = FFT2( sum_g psi1_g .* IFFT2( (fa*Fhat + fb) .* FFT2(pad(phi_g .* src)) ) )
// (I1Args as k_i1acc).  The output slot holds the physical-domain sum until the final FFT2.
template <typename R>
__global__ __launch_bounds__(big::TB<R>) void k_i1acc_big(I1Args a)
{
	using namespace big;
	constexpr int T = TB<R>, L = T / G;         // L: lines per sweep
	__shared__ cx<R> tw[TW];
	load_twiddles_g<R, P, G, T>(tw, a.tw256);
	const int lb = xcd_block();
	const int chunk = lb % a.nchunk;
	const int inst = (lb / a.nchunk) % a.ninst;
	const int cell = lb / (a.nchunk * a.ninst);
	const int qa = (int)(((long long)chunk * a.nterm) / a.nchunk);
	const int qb = (int)(((long long)(chunk + 1) * a.nterm) / a.nchunk);
	const int lt = threadIdx.x / G, p = threadIdx.x % G;
	const double2* g = a.src + ((size_t)cell * a.nsrc + a.instSrc[inst]) * (N * N);
	const cx<R>* Fh = (const cx<R>*)a.Fhat + ((size_t)cell * a.nFh + a.instFh[inst]) * PP;
	cx<R>* xs = (cx<R>*)a.xs + (size_t)blockIdx.x * PP;
	cx<R>* acc = (cx<R>*)a.out + (((size_t)cell * a.ninst + inst) * a.nchunk + chunk) * PP;
	const R fa = (R)a.fa, fb = (R)a.fb;
	// every acc element is read and written by one thread only until the final transform
#pragma unroll 1
	for (int sw = 0; sw < P / L; sw++) {
		const int r = sw * L + lt;
#pragma unroll
		for (int s = 0; s < E; s++) acc[r * P + p + G * s] = czero<R>();
	}
	__syncthreads();
	for (int q = qa; q < qb; q++) {
		pad_rows<R, L>(g, a.phi + (size_t)q * (N * N), 1, 0, xs, tw, lt, p);
		__syncthreads();
		// columns: FFT along l, times (fa Fhat + fb), IFFT along l
#pragma unroll 1
		for (int sw = 0; sw < P / L; sw++) {
			const int c = sw * L + lt;
			cx<R> v[E];
			load_padded_col<R>(v, xs, 0, c, p);
			line_fwd<R, P, G, E>(v, p, tw);
			const int r0 = blk(p);
#pragma unroll
			for (int u = 0; u < E; u++) {
				cx<R> f = Fh[(r0 + u) * P + c];
				f = mk<R>(fa * f.x + fb, fa * f.y);
				v[u] = cmul<R>(f, v[u]);
			}
			line_bwd<R, P, G, E>(v, p, tw);
#pragma unroll
			for (int s = 0; s < E; s++) xs[(p + G * s) * P + c] = v[s];
		}
		__syncthreads();
		// rows: IFFT along k, times psi1_g, into the sum
		const R* ps = (const R*)a.psi1n + (size_t)q * PP;
#pragma unroll 1
		for (int sw = 0; sw < P / L; sw++) {
			const int r = sw * L + lt;
			cx<R> v[E];
			const cx<R>* row = xs + r * P + blk(p);
#pragma unroll
			for (int u = 0; u < E; u++) v[u] = row[u];
			line_bwd<R, P, G, E>(v, p, tw);
#pragma unroll
			for (int s = 0; s < E; s++) {
				const int i = r * P + p + G * s;
				const R w = ps[i];
				cx<R> t = acc[i];
				t.x += w * v[s].x;
				t.y += w * v[s].y;
				acc[i] = t;
			}
		}
		__syncthreads();
	}
	// FFT2 of the sum in place: rows, then columns (each line is read whole before it is written)
#pragma unroll 1
	for (int sw = 0; sw < P / L; sw++) {
		const int r = sw * L + lt;
		cx<R> v[E];
#pragma unroll
		for (int s = 0; s < E; s++) v[s] = acc[r * P + p + G * s];
		line_fwd<R, P, G, E>(v, p, tw);
		cx<R>* row = acc + r * P + blk(p);
#pragma unroll
		for (int u = 0; u < E; u++) row[u] = v[u];
	}
	__syncthreads();
#pragma unroll 1
	for (int sw = 0; sw < P / L; sw++) {
		const int c = sw * L + lt;
		cx<R> v[E];
#pragma unroll
		for (int s = 0; s < E; s++) v[s] = acc[(p + G * s) * P + c];
		line_fwd<R, P, G, E>(v, p, tw);
		const int r0 = blk(p);
#pragma unroll
		for (int u = 0; u < E; u++) acc[(r0 + u) * P + c] = v[u];
	}
}

// ---------------------------------------------------------------------------------------------
// out[cell][inst][chunk] = sum_q FFT2(pad(psi1_q .* f)) .* FFT2(pad(psi2_q .* g))  (I2Args as
// k_i2acc, complex path).  Both pass-A grids share the scratch slot: f's rows at the padded
// positions, g's rows at the empty ones (row ^ 128).
template <typename R>
__global__ __launch_bounds__(big::TB<R>) void k_i2acc_big(I2Args a)
{
	using namespace big;
	constexpr int T = TB<R>, L = T / G;         // L: lines per sweep
	__shared__ cx<R> tw[TW];
	load_twiddles_g<R, P, G, T>(tw, a.tw256);
	const int lb = xcd_block();
	const int chunk = lb % a.nchunk;
	const int inst = (lb / a.nchunk) % a.ninst;
	const int cell = lb / (a.nchunk * a.ninst);
	const int qa = (int)(((long long)chunk * a.nterm) / a.nchunk);
	const int qb = (int)(((long long)(chunk + 1) * a.nterm) / a.nchunk);
	const int lt = threadIdx.x / G, p = threadIdx.x % G;
	const double2* f = a.src + ((size_t)cell * a.nsrc + a.instF[inst]) * (N * N);
	const double2* g = a.src + ((size_t)cell * a.nsrc + a.instG[inst]) * (N * N);
	cx<R>* xs = (cx<R>*)a.xs + (size_t)blockIdx.x * PP;
	cx<R>* acc = (cx<R>*)a.out + (((size_t)cell * a.ninst + inst) * a.nchunk + chunk) * PP;
	const int r0 = blk(p);
	// every output element is read and written by one thread only (columns, blocked)
#pragma unroll 1
	for (int sw = 0; sw < P / L; sw++) {
		const int c = sw * L + lt;
#pragma unroll
		for (int u = 0; u < E; u++) acc[(r0 + u) * P + c] = czero<R>();
	}
	__syncthreads();
	for (int q = qa; q < qb; q++) {
		pad_rows<R, L>(f, a.psi1 + (size_t)q * (N * N), 1, 0, xs, tw, lt, p);
		pad_rows<R, L>(g, a.psi2 + (size_t)q * (N * N), 1, 128, xs, tw, lt, p);
		__syncthreads();
#pragma unroll 1
		for (int sw = 0; sw < P / L; sw++) {
			const int c = sw * L + lt;
			cx<R> x[E], y[E];
			load_padded_col<R>(x, xs, 0, c, p);
			line_fwd<R, P, G, E>(x, p, tw);
			load_padded_col<R>(y, xs, 128, c, p);
			line_fwd<R, P, G, E>(y, p, tw);
#pragma unroll
			for (int u = 0; u < E; u++) {
				const int i = (r0 + u) * P + c;
				acc[i] = cadd<R>(acc[i], cmul<R>(x[u], y[u]));
			}
		}
		__syncthreads();
	}
}

// ---------------------------------------------------------------------------------------------
// acc[cell][out][chunk] (+)= sum_q sum_inst w F[f] .* K_q[Shat[g] .* F[h]],
// K_q[X] = FFT2(psi2_q .* IFFT2(X))  (I3Args as k_i3acc, complex path).  All instances of the
// output (at most 4) go through each pass together, in NI = 4 scratch grids per workgroup: rows
// (IFFT along k of S .* H), columns (IFFT along l, times psi2_q read once, FFT along l), rows (FFT
// along k, times w F, summed over the instances and added to the accumulator slot once per q).
// PAIR: packed pairs of real grids, as k_i3acc with PAIR (H one component of a packed F grid,
// instH = 2 slot + component; two weights per instance; the accumulator's real part only).
template <typename R, bool PAIR>
__global__ __launch_bounds__(big::TB<R>) void k_i3acc_big(I3Args a)
{
	using namespace big;
	constexpr int T = TB<R>, L = T / G;         // L: lines per sweep
	__shared__ cx<R> tw[TW];
	load_twiddles_g<R, P, G, T>(tw, a.tw256);
	const int lb = xcd_block();
	const int chunk = lb % a.nchunk;
	const int out = (lb / a.nchunk) % a.nout;
	const int cell = lb / (a.nchunk * a.nout);
	const int qa = (int)(((long long)chunk * a.nq) / a.nchunk);
	const int qb = (int)(((long long)(chunk + 1) * a.nq) / a.nchunk);
	const int lt = threadIdx.x / G, p = threadIdx.x % G;
	cx<R>* xs = (cx<R>*)a.xs + (size_t)blockIdx.x * (I3_NI * PP);
	cx<R>* accp = (cx<R>*)a.acc + (((size_t)cell * a.nout + out) * a.nchunk + chunk) * PP;
	const cx<R>* Sh = (const cx<R>*)a.Shat + (size_t)cell * a.nS * PP;
	const int ni = a.ninst[out];
	const int r0 = blk(p);
	// every accumulator element is read and written by one thread only (rows, blocked)
	if (!a.accumulate)
#pragma unroll 1
		for (int sw = 0; sw < P / L; sw++) {
			cx<R>* row = accp + (sw * L + lt) * P + r0;
#pragma unroll
			for (int u = 0; u < E; u++) row[u] = czero<R>();
		}
	__syncthreads();
#pragma unroll 1
	for (int qq = qa; qq < qb; qq++) {
		const cx<R>* Fq = (const cx<R>*)a.F + ((size_t)cell * a.nq + qq) * a.nF * PP;
		const R* ps = (const R*)a.psi2n + (size_t)(a.q0 + qq) * PP;
#pragma unroll 1
		for (int sw = 0; sw < P / L; sw++) {
			const int rb = (sw * L + lt) * P;
#pragma unroll 1
			for (int in = 0; in < ni; in++) {
				const int hs = a.instH[out][in];
				const cx<R>* Sg = uni(Sh, (long long)a.instS[out][in] * PP) + rb + r0;
				const cx<R>* Fh = uni(Fq, (long long)(PAIR ? hs >> 1 : hs) * PP) + rb + r0;
				cx<R> v[E];
				if constexpr (PAIR) {
					const bool hy = hs & 1;
#pragma unroll
					for (int u = 0; u < E; u++) {
						const cx<R> hh = Fh[u];
						const R hv = hy ? hh.y : hh.x;
						const cx<R> sv = Sg[u];
						v[u] = mk<R>(sv.x * hv, sv.y * hv);
					}
				} else {
#pragma unroll
					for (int u = 0; u < E; u++) v[u] = cmul<R>(Sg[u], Fh[u]);
				}
				line_bwd<R, P, G, E>(v, p, tw);
				cx<R>* x = uni(xs, (long long)in * PP) + rb + p;
#pragma unroll
				for (int s = 0; s < E; s++) x[G * s] = v[s];
			}
		}
		__syncthreads();
#pragma unroll 1
		for (int sw = 0; sw < P / L; sw++) {
			const int c = sw * L + lt;
			R m[E];
#pragma unroll
			for (int s = 0; s < E; s++) m[s] = ps[(p + G * s) * P + c];
#pragma unroll 1
			for (int in = 0; in < ni; in++) {
				cx<R>* x = uni(xs, (long long)in * PP) + r0 * P + c;
				cx<R> v[E];
#pragma unroll
				for (int u = 0; u < E; u++) v[u] = x[u * P];
				line_bwd<R, P, G, E>(v, p, tw);
#pragma unroll
				for (int s = 0; s < E; s++) {
					v[s].x *= m[s];
					v[s].y *= m[s];
				}
				line_fwd<R, P, G, E>(v, p, tw);
#pragma unroll
				for (int u = 0; u < E; u++) x[u * P] = v[u];
			}
		}
		__syncthreads();
#pragma unroll 1
		for (int sw = 0; sw < P / L; sw++) {
			const int rb = (sw * L + lt) * P;
			cx<R> sum[E];
#pragma unroll
			for (int u = 0; u < E; u++) sum[u] = czero<R>();
#pragma unroll 1
			for (int in = 0; in < ni; in++) {
				const cx<R>* x = uni((const cx<R>*)xs, (long long)in * PP) + rb + p;
				cx<R> v[E];
#pragma unroll
				for (int s = 0; s < E; s++) v[s] = x[G * s];
				line_fwd<R, P, G, E>(v, p, tw);
				const cx<R>* Ff = uni(Fq, (long long)a.instF[out][in] * PP) + rb + r0;
				if constexpr (PAIR) {
					const R wa = (R)a.instW[out][2 * in], wb = (R)a.instW[out][2 * in + 1];
#pragma unroll
					for (int u = 0; u < E; u++) {
						const cx<R> f = Ff[u];
						sum[u].x += (wa * f.x) * v[u].x + (wb * f.y) * v[u].y;
					}
				} else {
					const R w = (R)a.instW[out][in];
#pragma unroll
					for (int u = 0; u < E; u++) {
						const cx<R> y = cmul<R>(Ff[u], v[u]);
						sum[u].x += w * y.x;
						sum[u].y += w * y.y;
					}
				}
			}
			cx<R>* row = accp + rb + r0;
#pragma unroll
			for (int u = 0; u < E; u++) row[u] = cadd<R>(row[u], sum[u]);
		}
		__syncthreads();
	}
}

// ---------------------------------------------------------------------------------------------
// Combination, IFFT2, Extract and epilogue (CombineArgs as k_combine)
template <typename R>
__global__ __launch_bounds__(big::TB<R>) void k_combine_big(CombineArgs a)
{
	using namespace big;
	constexpr int T = TB<R>, L = T / G;         // L: lines per sweep
	__shared__ cx<R> tw[TW];
	load_twiddles_g<R, P, G, T>(tw, a.tw256);
	const int o = blockIdx.x % a.nout;
	const int cell = blockIdx.x / a.nout;
	const int lt = threadIdx.x / G, p = threadIdx.x % G;
	cx<R>* xs = (cx<R>*)a.xs + (size_t)blockIdx.x * PP;
	const int r0 = blk(p);
	__syncthreads();
#pragma unroll 1
	for (int sw = 0; sw < P / L; sw++) {
		const int rb = (sw * L + lt) * P;
		cx<R> v[E];
#pragma unroll
		for (int u = 0; u < E; u++) v[u] = czero<R>();
		for (int it = 0; it < a.nterm[o]; it++) {
			const CTerm c = a.terms[o * a.maxTerms + it];
			const cx<R>* Xp = (const cx<R>*)c.X + (size_t)cell * c.xcell + rb + r0;
			const cx<R>* Yp = (const cx<R>*)c.Y + (size_t)cell * c.ycell + rb + r0;
			const R ax = (R)c.ax, bx = (R)c.bx, cf = (R)c.coef;
#pragma unroll
			for (int u = 0; u < E; u++) {
				cx<R> y = mk<R>(R(1), R(0));
				if (c.Y) {
					y = Yp[u];
					for (int n = 1; n < c.nsum; n++) y = cadd<R>(y, Yp[(size_t)n * c.ysum + u]);
				}
				if (c.X) {
					cx<R> x = Xp[u];
					x = mk<R>(ax * x.x + bx, ax * x.y);
					y = cmul<R>(x, y);
				}
				v[u].x += cf * y.x;
				v[u].y += cf * y.y;
			}
		}
		line_bwd<R, P, G, E>(v, p, tw);
#pragma unroll
		for (int s = 0; s < E; s++) xs[rb + p + G * s] = v[s];
	}
	__syncthreads();
	// the N columns that the Extract keeps (src/grid.c:195-210)
	const int comp = a.outComp[o];
	const size_t cb = (size_t)cell * 4 * (N * N);
#pragma unroll 1
	for (int sw = 0; sw < N / L; sw++) {
		const int kn = sw * L + lt;
		const int c = padrow(kn);
		cx<R> v[E];
#pragma unroll
		for (int u = 0; u < E; u++) v[u] = xs[(r0 + u) * P + c];
		line_bwd<R, P, G, E>(v, p, tw);
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
}

} // namespace qbf
