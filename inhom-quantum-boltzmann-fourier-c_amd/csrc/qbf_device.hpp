// qbf_device.hpp -- device building blocks for batched 2-D FFT pipelines on gfx950.
//
// A P x P complex grid (row index l, column index k, element (l, k)) is owned by
// one workgroup of T threads.  G = T / P threads share one "line" (a row or a
// column); each thread holds E = P / G elements of its line in registers.
// Four register layouts are used:
//   R_i  line = row l,    thread p holds k = p + G*s          (s < E)  interleaved
//   R_b  line = row l,    thread p holds k = E*brev_G(p) + u  (u < E)  blocked
//   C_i  line = column k, thread p holds l = p + G*s
//   C_b  line = column k, thread p holds l = E*brev_G(p) + u
// A forward line FFT maps an interleaved layout to the blocked one (in-register
// DFT of length E, twiddle, radix-2 butterflies across the G lanes of the line);
// the backward line FFT maps blocked to interleaved.  Rows and columns are
// exchanged through LDS (R_i -> C_b and C_b -> R_i only).  Grids that live in
// global memory between kernels are stored in "native" order: element u of
// thread t at offset u*T + t, so every load and store of a wave is coalesced.
//
// FFT sign convention: forward = exp(-2 pi i jk/n), backward = exp(+2 pi i jk/n),
// both unnormalised (FFTW_FORWARD / FFTW_BACKWARD, as used by the reference
// src/integrals.c:48-51).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qbf {

template <typename R> struct cplx_sel;
template <> struct cplx_sel<float> { using type = float2; };
template <> struct cplx_sel<double> { using type = double2; };
template <typename R> using cx = typename cplx_sel<R>::type;

template <typename R>
__device__ __forceinline__ cx<R> mk(R a, R b)
{
	cx<R> r;
	r.x = a;
	r.y = b;
	return r;
}
template <typename R> __device__ __forceinline__ cx<R> cadd(cx<R> a, cx<R> b) { return mk<R>(a.x + b.x, a.y + b.y); }
template <typename R> __device__ __forceinline__ cx<R> csub(cx<R> a, cx<R> b) { return mk<R>(a.x - b.x, a.y - b.y); }
template <typename R> __device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b)
{
	return mk<R>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename R> __device__ __forceinline__ cx<R> cmulc(cx<R> a, cx<R> b)
{
	return mk<R>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename R> __device__ __forceinline__ cx<R> cscale(cx<R> a, R s) { return mk<R>(a.x * s, a.y * s); }
template <typename R> __device__ __forceinline__ cx<R> czero() { return mk<R>(R(0), R(0)); }

// cos / sin (2 pi j / 64), j < 64, correctly rounded double literals.
struct Tw64 {
	static constexpr double c[64] = {
		1.0, 0.9951847266721969, 0.9807852804032304, 0.9569403357322088,
		0.9238795325112867, 0.881921264348355, 0.8314696123025452, 0.773010453362737,
		0.7071067811865476, 0.6343932841636455, 0.5555702330196023, 0.4713967368259978,
		0.38268343236508984, 0.29028467725446233, 0.19509032201612833, 0.09801714032956077,
		6.123233995736766e-17, -0.09801714032956065, -0.1950903220161282, -0.29028467725446216,
		-0.3826834323650897, -0.4713967368259977, -0.555570233019602, -0.6343932841636454,
		-0.7071067811865475, -0.773010453362737, -0.8314696123025453, -0.8819212643483549,
		-0.9238795325112867, -0.9569403357322088, -0.9807852804032304, -0.9951847266721968,
		-1.0, -0.9951847266721969, -0.9807852804032304, -0.9569403357322089,
		-0.9238795325112868, -0.881921264348355, -0.8314696123025455, -0.7730104533627371,
		-0.7071067811865477, -0.6343932841636459, -0.5555702330196022, -0.47139673682599786,
		-0.38268343236509034, -0.29028467725446244, -0.19509032201612866, -0.09801714032956045,
		-1.8369701987210297e-16, 0.09801714032956009, 0.1950903220161283, 0.29028467725446205,
		0.38268343236509, 0.4713967368259976, 0.5555702330196018, 0.6343932841636456,
		0.7071067811865474, 0.7730104533627367, 0.8314696123025452, 0.8819212643483548,
		0.9238795325112865, 0.9569403357322088, 0.9807852804032303, 0.9951847266721969};
	static constexpr double sn[64] = {
		0.0, 0.0980171403295606, 0.19509032201612825, 0.29028467725446233,
		0.3826834323650898, 0.47139673682599764, 0.5555702330196022, 0.6343932841636455,
		0.7071067811865475, 0.773010453362737, 0.8314696123025452, 0.8819212643483549,
		0.9238795325112867, 0.9569403357322089, 0.9807852804032304, 0.9951847266721968,
		1.0, 0.9951847266721969, 0.9807852804032304, 0.9569403357322089,
		0.9238795325112867, 0.881921264348355, 0.8314696123025455, 0.7730104533627371,
		0.7071067811865476, 0.6343932841636455, 0.5555702330196022, 0.47139673682599786,
		0.3826834323650899, 0.2902846772544624, 0.1950903220161286, 0.09801714032956083,
		1.2246467991473532e-16, -0.09801714032956059, -0.19509032201612836, -0.2902846772544621,
		-0.38268343236508967, -0.47139673682599764, -0.555570233019602, -0.6343932841636453,
		-0.7071067811865475, -0.7730104533627367, -0.8314696123025452, -0.8819212643483549,
		-0.9238795325112865, -0.9569403357322088, -0.9807852804032303, -0.9951847266721969,
		-1.0, -0.9951847266721969, -0.9807852804032304, -0.9569403357322089,
		-0.9238795325112866, -0.881921264348355, -0.8314696123025455, -0.7730104533627369,
		-0.7071067811865477, -0.6343932841636459, -0.5555702330196022, -0.4713967368259979,
		-0.3826834323650904, -0.2902846772544625, -0.19509032201612872, -0.0980171403295605};
	static constexpr double s(int j) { return sn[j]; }
};

template <int E>
__host__ __device__ constexpr int brev(int i)
{
	int r = 0;
	for (int b = 1, rb = E >> 1; b < E; b <<= 1, rb >>= 1)
		if (i & b) r |= rb;
	return r;
}

template <int X>
__host__ __device__ constexpr int ilog2c()
{
	int l = 0;
	while ((1 << l) < X) l++;
	return l;
}

// c * exp(SIGN * 2 pi i j / m) for compile-time (after unrolling) j, m <= 64.
template <typename R, int SIGN>
__device__ __forceinline__ cx<R> twmul(cx<R> c, int j, int m)
{
	if (j == 0) return c;
	if (4 * j == m) return SIGN < 0 ? mk<R>(c.y, -c.x) : mk<R>(-c.y, c.x);
	if (4 * j == 3 * m) return SIGN < 0 ? mk<R>(-c.y, c.x) : mk<R>(c.y, -c.x);
	if (2 * j == m) return mk<R>(-c.x, -c.y);
	const R cr = (R)Tw64::c[j * (64 / m)];
	const R si = (R)(SIGN * Tw64::s(j * (64 / m)));
	return mk<R>(c.x * cr - c.y * si, c.x * si + c.y * cr);
}

// In-register DFT of length E (power of two <= 64), natural order in and out.
template <typename R, int E, int SIGN>
__device__ __forceinline__ void dft(cx<R>* v)
{
	if constexpr (E > 1) {
		cx<R> w[E];
#pragma unroll
		for (int i = 0; i < E; i++) w[i] = v[brev<E>(i)];
#pragma unroll
		for (int m = 2; m <= E; m <<= 1) {
#pragma unroll
			for (int b = 0; b < E; b += m) {
#pragma unroll
				for (int j = 0; j < m / 2; j++) {
					const cx<R> t = twmul<R, SIGN>(w[b + j + m / 2], j, m);
					const cx<R> a = w[b + j];
					w[b + j] = cadd<R>(a, t);
					w[b + j + m / 2] = csub<R>(a, t);
				}
			}
		}
#pragma unroll
		for (int i = 0; i < E; i++) v[i] = w[i];
	}
}

// Exchange a 32-bit value with lane (lane ^ m).  m is a compile-time constant after
// unrolling: lane distances 1 and 2 use DPP quad permutes (VALU, no LDS traffic),
// 4, 8, 16 the ds_swizzle xor mode, 32 a bpermute.
__device__ __forceinline__ int xor_lane_i32(int x, int m)
{
	switch (m) {
	case 1: return __builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
	case 2: return __builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
	case 4: return __builtin_amdgcn_ds_swizzle(x, 0x101F);               // and 0x1f, xor 4
	case 8: return __builtin_amdgcn_ds_swizzle(x, 0x201F);
	case 16: return __builtin_amdgcn_ds_swizzle(x, 0x401F);
	default: return __shfl_xor(x, m);
	}
}
__device__ __forceinline__ float xor_lane(float x, int m) { return __int_as_float(xor_lane_i32(__float_as_int(x), m)); }
__device__ __forceinline__ double xor_lane(double x, int m)
{
	const long long b = __double_as_longlong(x);
	const int lo = xor_lane_i32((int)(b & 0xffffffffLL), m);
	const int hi = xor_lane_i32((int)(b >> 32), m);
	return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

template <typename R>
__device__ __forceinline__ cx<R> shfl_xor_c(cx<R> v, int m)
{
	return mk<R>(xor_lane(v.x, m), xor_lane(v.y, m));
}

// Geometry of one P x P grid per workgroup of T threads.
template <int P_, int T_>
struct Geo {
	static constexpr int P = P_;
	static constexpr int T = T_;
	static constexpr int G = T_ / P_;       // threads per line
	static constexpr int E = P_ / G;        // elements per thread
	static constexpr int PITCH = P_ + G;    // LDS row pitch (elements), see Xpose
	static_assert(G * P_ == T_, "T must be a multiple of P");
	static_assert(G * E == P_, "G must divide P");
	static_assert(G <= 64 && (64 % G) == 0, "a line must sit inside one wavefront");
};

// Logical block index that keeps consecutive logical blocks on one XCD.  The dispatcher hands
// block b to XCD b % 8, so XCD x runs the logical blocks [x n/8, (x+1) n/8): workgroups that read
// the same grids (the outputs of one cell) share that XCD's L2.  A permutation of [0, n) for any
// n divisible by 8 (identity otherwise), so correctness never depends on the dispatch order.
__device__ __forceinline__ int xcd_block()
{
	const int n = gridDim.x, b = blockIdx.x;
	return (n & 7) ? b : (b & 7) * (n >> 3) + (b >> 3);
}

// An SGPR zero the compiler cannot see through: adding it to an LDS index keeps table loads
// next to their use instead of hoisting them (and their registers) out of the term loops.
__device__ __forceinline__ int opaque_zero()
{
	int z = 0;
	asm volatile("" : "+s"(z));
	return z;
}

// p + off with an opaque SGPR term, so the address (and the load) is formed next to its
// use; p keeps its address space (a global pointer stays a global load).
template <typename X>
__device__ __forceinline__ X* uni(X* p, long long off)
{
	typedef __attribute__((address_space(1))) X gX;
	unsigned long long a = (unsigned long long)(p + off);
	asm volatile("" : "+s"(a));
	return (X*)(gX*)a;
}

// Materialise a value in VGPRs here: keeps a load unconditional (not sunk into a branch
// of the lanes that use it).
template <typename R>
__device__ __forceinline__ cx<R> pin(cx<R> w)
{
	asm volatile("" : "+v"(w.x), "+v"(w.y));
	return w;
}

// Twiddle tables of a P-point line FFT over T threads (G = T / P lanes per line): w^j (P entries),
// the per-lane table (P entries) and the cross-lane stage twiddles (G entries).
template <int P, int T>
struct TwTab {
	static constexpr int SIZE = 2 * P + T / P;
};

// Layout of the per-lane twiddle table w^(p u): float grids interleave it by lane, [u][p], so
// that the G distinct addresses of one read fall into distinct banks; double grids keep it
// [p][u] (their N = 64 kernels are register-bound, and the interleaved addressing costs them
// registers).
template <typename R, int G, int E>
__device__ __forceinline__ constexpr int tw_lane(int p) { return sizeof(R) == 4 ? p : p * E; }
template <typename R, int G>
__device__ __forceinline__ constexpr int tw_step() { return sizeof(R) == 4 ? G : 1; }

// Forward FFT of one line: interleaved layout in, blocked layout out.
// twP[j] = (cos, sin)(2 pi j / P), j < P; the per-lane table (cos, sin)(2 pi p u / P) follows it
// (in LDS, laid out as tw_lane / tw_step below).
template <typename R, int P, int G, int E>
__device__ __forceinline__ void line_fwd(cx<R>* v, int p, const cx<R>* twP)
{
	dft<R, E, -1>(v);
	if constexpr (G > 1) {
		const cx<R>* tl = twP + P + tw_lane<R, G, E>(p) + opaque_zero();
#pragma unroll
		for (int u = 1; u < E; u++) v[u] = cmulc<R>(v[u], tl[tw_step<R, G>() * u]);
#pragma unroll
		for (int h = G / 2; h >= 1; h >>= 1) {
			// DIF butterfly across lanes p, p^h:  lo: a + b,  hi: (a - b) * conj(w)
			// as  (partner + sgn * self) * wl  with lane constants sgn = +-1, wl = 1 or conj(w)
			const bool hi = (p & h) != 0;
			const R sgn = hi ? R(-1) : R(1);
			if (h == 1) {                      // twiddle exp(0) = 1
#pragma unroll
				for (int u = 0; u < E; u++) {
					const cx<R> pv = shfl_xor_c<R>(v[u], h);
					v[u] = mk<R>(pv.x + sgn * v[u].x, pv.y + sgn * v[u].y);
				}
			} else if (h == 2) {               // twiddle w^((p & 1) P / 4) = 1 or i
				// the hi lanes with odd p multiply by conj(i) = -i: a swap and a sign (two selects)
				const bool q = hi && (p & 1);
#pragma unroll
				for (int u = 0; u < E; u++) {
					const cx<R> pv = shfl_xor_c<R>(v[u], h);
					const cx<R> t = mk<R>(pv.x + sgn * v[u].x, pv.y + sgn * v[u].y);
					v[u] = mk<R>(q ? t.y : t.x, q ? -t.x : t.y);
				}
			} else {
				const cx<R> w = pin<R>(twP[2 * P + h + (p & (h - 1)) + opaque_zero()]);
				const cx<R> wl = hi ? mk<R>(w.x, -w.y) : mk<R>(R(1), R(0));
#pragma unroll
				for (int u = 0; u < E; u++) {
					const cx<R> pv = shfl_xor_c<R>(v[u], h);
					const cx<R> t = mk<R>(pv.x + sgn * v[u].x, pv.y + sgn * v[u].y);
					v[u] = cmul<R>(t, wl);
				}
			}
		}
	}
}

// Backward FFT of one line: blocked layout in, interleaved layout out.
template <typename R, int P, int G, int E>
__device__ __forceinline__ void line_bwd(cx<R>* v, int p, const cx<R>* twP)
{
	if constexpr (G > 1) {
#pragma unroll
		for (int h = 1; h < G; h <<= 1) {
			// DIT butterfly across lanes p, p^h:  lo: a + w b,  hi: a - w b  (a on lo, b on hi):
			// the hi lane pre-multiplies its own value by w, then  partner' + sgn * self'
			const bool hi = (p & h) != 0;
			const R sgn = hi ? R(-1) : R(1);
			if (h == 1) {                      // twiddle exp(0) = 1
#pragma unroll
				for (int u = 0; u < E; u++) {
					const cx<R> pv = shfl_xor_c<R>(v[u], h);
					v[u] = mk<R>(pv.x + sgn * v[u].x, pv.y + sgn * v[u].y);
				}
			} else if (h == 2) {               // twiddle w^((p & 1) P / 4) = 1 or i
				// the hi lanes with odd p pre-multiply by i: a swap and a sign (two selects)
				const bool q = hi && (p & 1);
#pragma unroll
				for (int u = 0; u < E; u++) {
					const cx<R> sv = mk<R>(q ? -v[u].y : v[u].x, q ? v[u].x : v[u].y);
					const cx<R> pv = shfl_xor_c<R>(sv, h);
					v[u] = mk<R>(pv.x + sgn * sv.x, pv.y + sgn * sv.y);
				}
			} else {
				const cx<R> w = pin<R>(twP[2 * P + h + (p & (h - 1)) + opaque_zero()]);
				const cx<R> wl = hi ? w : mk<R>(R(1), R(0));
#pragma unroll
				for (int u = 0; u < E; u++) {
					const cx<R> sv = cmul<R>(v[u], wl);
					const cx<R> pv = shfl_xor_c<R>(sv, h);
					v[u] = mk<R>(pv.x + sgn * sv.x, pv.y + sgn * sv.y);
				}
			}
		}
		const cx<R>* tl = twP + P + tw_lane<R, G, E>(p) + opaque_zero();
#pragma unroll
		for (int u = 1; u < E; u++) v[u] = cmul<R>(v[u], tl[tw_step<R, G>() * u]);
	}
	dft<R, E, +1>(v);
}

// LDS exchanges.  The grid is stored row-major with pitch P + G and a skew of SKEW elements
// per block of E rows.  An 8-byte LDS access of a wave is served 128 bytes per phase: LANES
// consecutive lanes of one element each, conflict-free when their addresses are distinct modulo
// LANES elements.  (Measured with SQ_LDS_BANK_CONFLICT of k_i3acc for the column skews of the
// 512-, 256- and 128-byte phase models: 1.31e7, 4.72e6 and 5.2e5 cycles; lane-linear and stride-2
// probes do not tell the models apart.)  Writes / reads of rows (R_i) hit LANES/G consecutive
// rows times G consecutive columns: distinct because the pitch is G modulo LANES.  Writes /
// reads of columns (C_b) hit LANES/G consecutive columns in G row blocks E rows apart: E*pitch
// is a multiple of LANES, and the per-block skew LANES/G separates the blocks.
template <typename R, class GE>
struct Xpose {
	static constexpr int LANES = 128 / (int)sizeof(cx<R>);
	static constexpr int SKEW = LANES / GE::G;
	static constexpr int SIZE = GE::P * GE::PITCH + GE::G * SKEW;
	cx<R>* g0;
	__device__ __forceinline__ cx<R>* gp() const { return g0 + opaque_zero(); }
	__device__ __forceinline__ int a(int row, int col) const { return row * GE::PITCH + col + (row / GE::E) * SKEW; }
	// R_i (line = row) -> LDS
	__device__ __forceinline__ void put_Ri(const cx<R>* v, int line, int p) const
	{
		cx<R>* g = gp();
#pragma unroll
		for (int s = 0; s < GE::E; s++) g[a(line, p + GE::G * s)] = v[s];
	}
	// LDS -> C_b (line = column)
	__device__ __forceinline__ void get_Cb(cx<R>* v, int line, int p) const
	{
		const cx<R>* g = gp();
		const int r0 = GE::E * brev<GE::G>(p);
#pragma unroll
		for (int u = 0; u < GE::E; u++) v[u] = g[a(r0 + u, line)];
	}
	// C_b (line = column) -> LDS
	__device__ __forceinline__ void put_Cb(const cx<R>* v, int line, int p) const
	{
		cx<R>* g = gp();
		const int r0 = GE::E * brev<GE::G>(p);
#pragma unroll
		for (int u = 0; u < GE::E; u++) g[a(r0 + u, line)] = v[u];
	}
	// one element of row `row`
	__device__ __forceinline__ cx<R> get(int row, int col) const { return gp()[a(row, col)]; }
	// LDS -> R_i (line = row)
	__device__ __forceinline__ void get_Ri(cx<R>* v, int line, int p) const
	{
		const cx<R>* g = gp();
#pragma unroll
		for (int s = 0; s < GE::E; s++) v[s] = g[a(line, p + GE::G * s)];
	}
};

// Fill the LDS twiddle tables from a global table of (cos, sin)(2 pi j / 256) in double:
// twP[j] = w^j for j < P, the per-lane table w^(p*u) at twP[P + tw_lane(p) + tw_step * u], and the
// twiddles of the cross-lane stages at twP[2P + h + m] = w^(m P / 2h), m < h < G, kept apart so
// that the lanes of one stage read consecutive entries; w = exp(2 pi i / P).
// (load_twiddles_g: G lanes per line, filled by the NT threads of the workgroup)
template <typename R, int P, int G, int NT>
__device__ __forceinline__ void load_twiddles_g(cx<R>* twP, const double2* tw256)
{
	constexpr int E = P / G;
	for (int j = threadIdx.x; j < 2 * P + G; j += NT) {
		int k;
		if (j < P) {
			k = j;
		} else if (j < 2 * P) {
			k = sizeof(R) == 4 ? ((j - P) % G) * ((j - P) / G) : ((j - P) / E) * ((j - P) % E);
		} else {
			const int i = j - 2 * P;
			int h = 1;
			while (2 * h <= i) h *= 2;
			k = i == 0 ? 0 : (i - h) * (P / (2 * h));
		}
		const double2 w = tw256[k * (256 / P)];
		twP[j] = mk<R>((R)w.x, (R)w.y);
	}
}
template <typename R, int P, int T>
__device__ __forceinline__ void load_twiddles(cx<R>* twP, const double2* tw256)
{
	static_assert(TwTab<P, T>::SIZE == 2 * P + T / P, "table layout");
	load_twiddles_g<R, P, T / P, T>(twP, tw256);
}

// Native (coalesced) storage of a register-resident grid.
template <typename R, class GE>
__device__ __forceinline__ void store_native(cx<R>* dst, const cx<R>* v)
{
#pragma unroll
	for (int u = 0; u < GE::E; u++) uni(dst, (long long)u * GE::T)[threadIdx.x] = v[u];
}
template <typename R, class GE>
__device__ __forceinline__ void load_native(cx<R>* v, const cx<R>* src)
{
#pragma unroll
	for (int u = 0; u < GE::E; u++) v[u] = uni(src, (long long)u * GE::T)[threadIdx.x];
}

} // namespace qbf
