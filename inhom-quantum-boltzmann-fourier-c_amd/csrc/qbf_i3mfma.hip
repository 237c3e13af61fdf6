// qbf_i3mfma.hip -- the I3 K-operator accumulation (k_i3acc) at N = 64 (P = 128) with its line
// DFTs on the matrix cores: BF16 hi/lo split products accumulated in FP32.
//
// K_q[X] = FFT2(psi2_q .* IFFT2(X)) on a 128 x 128 complex grid (src/integrals.c:343-378: the I3
// term loop that this kernel's accumulation replaces, reassociated as in k_i3acc).  Every
// 128-point line DFT is factored 16 x 8:
//   backward (DIF), l = 8 n1 + n2, m = k1 + 16 k2:
//     Y[k1][n2] = sum_n1 w16^(n1 k1) x[8 n1 + n2]          DFT-16 GEMM on the matrix cores
//     Z[k1 + 16 k2] = sum_n2 w8^(n2 k2) w128^(n2 k1) Y      twiddle + DFT-8 on the FP32 VALU
//   forward (DIT), b = n1 + 16 n2, k = k2 + 8 k1: the same two stages in the opposite order,
// so the forward transform takes the backward one's output order as it stands.
// The DFT-16 GEMM: v_mfma_f32_16x16x32_bf16, the complex 16 x 16 matrix as a real 16 x 32 block
// per output part (real, imaginary), K = 16 complex inputs as (re, im) pairs, N = 16 grid lines.
// Operands are split x = hi + lo (bf16 each); a product is hi*hi + hi*lo + lo*hi (the lo*lo term,
// 2^-16 relative, is dropped), accumulated in FP32 by the matrix core.
//
// Work split: one workgroup of 8 waves per (cell, output, term range), as k_i3acc.  Wave w owns
// 16 grid columns in the first and last pass and 16 grid rows in the middle pass; the two
// transposes between them go through LDS (128 x 129 complex floats, conflict-free for all four
// access patterns).  Lane (i16, kq) = lane % 16, lane / 16 holds 32 elements:
//   columns pass  column kappa(w, i16), rows 32 kq + 8 r + t   (element d[t][r], r < 4, t < 8)
//   rows pass     row 16 w + i16,       columns 32 kq + 8 r + t
// The grids in global memory (spectra, F grids, accumulators) keep k_i3acc's R_b native order
// for T = 1024 threads: element (l, k) at (k % 16) * 1024 + l * 8 + brev8(k / 16); column
// kappa(w, i16) = 16 brev8(i16 % 8) + 2 w + i16 / 8 makes 8 lanes read 64 contiguous bytes.
#pragma once
#include "qbf_device.hpp"
#include "qbf_kernels.hpp"

namespace qbf {
namespace mf {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

constexpr int P = 128;
constexpr int NW = 8;                 // waves per workgroup
constexpr int T = 64 * NW;
constexpr int PITCH = P + 1;          // LDS row pitch (complex floats): odd, see above
constexpr int LDS_GRID = P * PITCH;

// LDS position of grid column k: (k % 16) * 8 + k / 16
__device__ __forceinline__ int colpos(int k) { return (k & 15) * 8 + (k >> 4); }

typedef __attribute__((ext_vector_type(2))) float f32x2;
typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2;

// hi = bf16(x) (round to nearest even), lo = bf16(x - hi), two values per conversion: the bf16 pair
// is read back as floats from its halves (bf16 is the upper half of a float)
__device__ __forceinline__ void split8(const float* x, bf16x8& hi, bf16x8& lo)
{
#pragma unroll
	for (int j = 0; j < 4; j++) {
		const f32x2 v = {x[2 * j], x[2 * j + 1]};
		const bf16x2 h = __builtin_convertvector(v, bf16x2);
		const unsigned hb = __builtin_bit_cast(unsigned, h);
		const f32x2 hf = {__uint_as_float(hb << 16), __uint_as_float(hb & 0xffff0000u)};
		const bf16x2 l = __builtin_convertvector(v - hf, bf16x2);
		hi[2 * j] = h[0];
		hi[2 * j + 1] = h[1];
		lo[2 * j] = l[0];
		lo[2 * j + 1] = l[1];
	}
}

// The DFT-16 matrix rows of this lane as A operands (lane (i16, kq): row k1 = i16, K slots
// 8 kq + 2 r + c for input n1 = 4 kq + r, c = 0 real / 1 imaginary part), for the real and the
// imaginary output part; SIGN = +1 backward, -1 forward.
struct Dft16 {
	bf16x8 re_hi, re_lo, im_hi, im_lo;
};
template <int SIGN>
__device__ __forceinline__ Dft16 dft16_rows(int i16, int kq, const double2* tw256)
{
	float are[8], aim[8];
#pragma unroll
	for (int r = 0; r < 4; r++) {
		const int n1 = 4 * kq + r;
		const int j = (n1 * i16) & 15;                  // angle 2 pi j / 16
		const double2 cs = tw256[16 * j];
		const float c = (float)cs.x, s = SIGN * (float)cs.y;
		// (a + i b) (c + i s) = (a c - b s) + i (a s + b c)
		are[2 * r] = c;
		are[2 * r + 1] = -s;
		aim[2 * r] = s;
		aim[2 * r + 1] = c;
	}
	Dft16 m;
	split8(are, m.re_hi, m.re_lo);
	split8(aim, m.im_hi, m.im_lo);
	return m;
}

// One DFT-16 GEMM over the column block held as d[0..3] (4 complex per lane, K = 16 complex
// inputs n1 = 4 kq + r); the result replaces d: output k1 = 4 kq + r.
__device__ __forceinline__ void dft16_mfma(float2* d, const Dft16& m)
{
	float x[8];
#pragma unroll
	for (int r = 0; r < 4; r++) {
		x[2 * r] = d[r].x;
		x[2 * r + 1] = d[r].y;
	}
	bf16x8 bh, bl;
	split8(x, bh, bl);
	const f32x4 z = {0.f, 0.f, 0.f, 0.f};
	f32x4 yr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(m.re_lo, bh, z, 0, 0, 0);
	yr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(m.re_hi, bl, yr, 0, 0, 0);
	yr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(m.re_hi, bh, yr, 0, 0, 0);
	f32x4 yi = __builtin_amdgcn_mfma_f32_16x16x32_bf16(m.im_lo, bh, z, 0, 0, 0);
	yi = __builtin_amdgcn_mfma_f32_16x16x32_bf16(m.im_hi, bl, yi, 0, 0, 0);
	yi = __builtin_amdgcn_mfma_f32_16x16x32_bf16(m.im_hi, bh, yi, 0, 0, 0);
#pragma unroll
	for (int r = 0; r < 4; r++) d[r] = make_float2(yr[r], yi[r]);
}

// Backward 128-point DFT of the 32 elements of a lane (DIF order in: d[t][r] = x[32 kq + 8 r + t];
// out: d[k2][r] = X[4 kq + r + 16 k2]).  tw[j] = w128^j = (cos, sin)(2 pi j / 128).
__device__ __forceinline__ void line_bwd128(float2 (*d)[4], const Dft16& mb, const float2* tw, int kq)
{
#pragma unroll
	for (int t = 0; t < 8; t++) dft16_mfma(d[t], mb);
#pragma unroll
	for (int r = 0; r < 4; r++) {
		const int k1 = 4 * kq + r;
		float2 col[8];
#pragma unroll
		for (int t = 0; t < 8; t++) col[t] = t == 0 ? d[0][r] : cmul<float>(d[t][r], tw[(t * k1) & (P - 1)]);
		dft<float, 8, +1>(col);
#pragma unroll
		for (int t = 0; t < 8; t++) d[t][r] = col[t];
	}
}

// Forward 128-point DFT (DIT order in: d[t][r] = x[4 kq + r + 16 t]; out: d[k2][r] = X[32 kq + 8 r + k2]).
__device__ __forceinline__ void line_fwd128(float2 (*d)[4], const Dft16& mf, const float2* tw, int kq)
{
#pragma unroll
	for (int r = 0; r < 4; r++) {
		const int n1 = 4 * kq + r;
		float2 col[8];
#pragma unroll
		for (int t = 0; t < 8; t++) col[t] = d[t][r];
		dft<float, 8, -1>(col);
#pragma unroll
		for (int t = 0; t < 8; t++) d[t][r] = t == 0 ? col[0] : cmulc<float>(col[t], tw[(t * n1) & (P - 1)]);
	}
#pragma unroll
	for (int t = 0; t < 8; t++) dft16_mfma(d[t], mf);
}

// element idx (the lane's index, 0 <= idx < P * P) from the uniform base p + off: the load takes
// the SGPR base + VGPR offset form
template <typename X>
__device__ __forceinline__ X& at(X* p, long long off, int idx)
{
	__builtin_assume(idx >= 0 && idx < P * P);
	return uni(p, off)[idx];
}

} // namespace mf

// acc[cell][out][chunk] (+)= sum_q sum_inst w F[f] .* K_q[Shat[g] .* F[h]], as k_i3acc<float, 64, 1024,
// false, PAIR> (same arguments and grid orders) except psi2n: [nqTotal][P*P] in this kernel's order
// (psi2 of row 16 w + i16, column 4 kq + r + 16 t at ((w * 8 + t) * 4 + r) * 64 + lane).
template <bool PAIR>
__global__ __launch_bounds__(mf::T) void k_i3acc_mfma(I3Args a)
{
	using namespace mf;
	__shared__ float2 lds[LDS_GRID];
	__shared__ float2 tw[P];
	for (int j = threadIdx.x; j < P; j += T) {
		const double2 w = a.tw256[2 * j];
		tw[j] = make_float2((float)w.x, (float)w.y);
	}
	__syncthreads();
	const int lb = xcd_block();
	const int chunk = lb % a.nchunk;
	const int out = (lb / a.nchunk) % a.nout;
	const int cell = lb / (a.nchunk * a.nout);
	const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63, i16 = lane & 15, kq = lane >> 4;
	const int qa = (int)(((long long)chunk * a.nq) / a.nchunk);
	const int qb = (int)(((long long)(chunk + 1) * a.nq) / a.nchunk);
	// columns pass: column kappa = 16 brev8(pb) + u; native offset of (l, kappa) = u * 1024 + 8 l + pb
	const int u = 2 * w + (i16 >> 3), pb = i16 & 7;
	// (formed from threadIdx.x directly, whose range the compiler knows: the loads then take the
	// SGPR base + VGPR offset form)
	const int tx = threadIdx.x;
	const int gb = (2 * (tx >> 6) + ((tx >> 3) & 1)) * 1024 + 256 * ((tx >> 4) & 3) + (tx & 7);   // + 64 r + 8 t: row 32 kq + 8 r + t
	const int cpos = u * 8 + brev<8>(pb);                    // LDS column position of kappa
	const int prow = (16 * w + i16) * PITCH;                 // rows pass: LDS row of this lane
	const Dft16 mb = dft16_rows<+1>(i16, kq, a.tw256), mfw = dft16_rows<-1>(i16, kq, a.tw256);
	float2* accp = (float2*)a.acc + (((size_t)cell * a.nout + out) * a.nchunk + chunk) * (P * P);
	float acc[8][4][PAIR ? 1 : 2];
	if (a.accumulate) {
#pragma unroll
		for (int r = 0; r < 4; r++)
#pragma unroll
			for (int t = 0; t < 8; t++) {
				const float2 v = at(accp, 64 * r + 8 * t, gb);
				acc[t][r][0] = v.x;
				if constexpr (!PAIR) acc[t][r][PAIR ? 0 : 1] = v.y;
			}
	} else {
#pragma unroll
		for (int t = 0; t < 8; t++)
#pragma unroll
			for (int r = 0; r < 4; r++)
#pragma unroll
				for (int c = 0; c < (PAIR ? 1 : 2); c++) acc[t][r][c] = 0.f;
	}
	const float2* Sh = (const float2*)a.Shat + (size_t)cell * a.nS * (P * P);
	const int ni = a.ninst[out];
	float2 d[8][4];
	for (int qq = qa; qq < qb; qq++) {
		const float2* Fq = (const float2*)a.F + ((size_t)cell * a.nq + qq) * a.nF * (P * P);
		const float* ps = (const float*)a.psi2n + (size_t)(a.q0 + qq) * (P * P) + (size_t)w * 2048;
		for (int in = 0; in < ni; in++) {
			const float2* Sg = Sh + (size_t)a.instS[out][in] * (P * P);
			const int hs = a.instH[out][in];
			const float2* Fh = Fq + (size_t)(PAIR ? hs >> 1 : hs) * (P * P);
			const float2* Ff = Fq + (size_t)a.instF[out][in] * (P * P);
			// X = S .* H in the columns-pass order
#pragma unroll
			for (int r = 0; r < 4; r++)
#pragma unroll
				for (int t = 0; t < 8; t++) {
					const float2 sv = at(Sg, 64 * r + 8 * t, gb);
					const float2 h = at(Fh, 64 * r + 8 * t, gb);
					if constexpr (PAIR) {
						const float hv = (hs & 1) ? h.y : h.x;
						d[t][r] = make_float2(sv.x * hv, sv.y * hv);
					} else {
						d[t][r] = cmul<float>(sv, h);
					}
				}
			// IFFT along l (columns pass), out: rows 4 kq + r + 16 t
			__builtin_amdgcn_sched_barrier(0);
			line_bwd128(d, mb, tw, kq);
			__syncthreads();                      // the previous application's reads of lds are done
#pragma unroll
			for (int t = 0; t < 8; t++)
#pragma unroll
				for (int r = 0; r < 4; r++) lds[(4 * kq + r + 16 * t) * PITCH + cpos] = d[t][r];
			__syncthreads();
			// rows pass: row 16 w + i16, columns 32 kq + 8 r + t
#pragma unroll
			for (int t = 0; t < 8; t++)
#pragma unroll
				for (int r = 0; r < 4; r++) d[t][r] = lds[prow + colpos(32 * kq + 8 * r + t)];
			line_bwd128(d, mb, tw, kq);           // IFFT along k, out: columns 4 kq + r + 16 t
			// keep the table and operand loads next to their use (hoisted, they would hold registers
			// across the line DFTs)
			__builtin_amdgcn_sched_barrier(0);
#pragma unroll
			for (int t = 0; t < 8; t++)
#pragma unroll
				for (int r = 0; r < 4; r++) {
					const float m = at(ps, (t * 4 + r) * 64, tx & 63);
					d[t][r].x *= m;
					d[t][r].y *= m;
				}
			__builtin_amdgcn_sched_barrier(0);
			line_fwd128(d, mfw, tw, kq);          // FFT along k, out: columns 32 kq + 8 r + t
			__syncthreads();                      // every wave has read its rows
#pragma unroll
			for (int t = 0; t < 8; t++)
#pragma unroll
				for (int r = 0; r < 4; r++) lds[prow + colpos(32 * kq + 8 * r + t)] = d[t][r];
			__syncthreads();
			// columns pass again: rows 4 kq + r + 16 t (the forward DIT input order)
#pragma unroll
			for (int t = 0; t < 8; t++)
#pragma unroll
				for (int r = 0; r < 4; r++) d[t][r] = lds[(4 * kq + r + 16 * t) * PITCH + cpos];
			line_fwd128(d, mfw, tw, kq);          // FFT along l, out: rows 32 kq + 8 r + t
			__builtin_amdgcn_sched_barrier(0);
			if constexpr (PAIR) {
				const float wa = (float)a.instW[out][2 * in], wb = (float)a.instW[out][2 * in + 1];
#pragma unroll
				for (int r = 0; r < 4; r++)
#pragma unroll
					for (int t = 0; t < 8; t++) {
						const float2 f = at(Ff, 64 * r + 8 * t, gb);
						acc[t][r][0] += (wa * f.x) * d[t][r].x + (wb * f.y) * d[t][r].y;
					}
			} else {
				const float wt = (float)a.instW[out][in];
#pragma unroll
				for (int r = 0; r < 4; r++)
#pragma unroll
					for (int t = 0; t < 8; t++) {
						const float2 y = cmul<float>(at(Ff, 64 * r + 8 * t, gb), d[t][r]);
						acc[t][r][0] += wt * y.x;
						acc[t][r][PAIR ? 0 : 1] += wt * y.y;
					}
			}
		}
	}
#pragma unroll
	for (int r = 0; r < 4; r++)
#pragma unroll
		for (int t = 0; t < 8; t++)
			at(accp, 64 * r + 8 * t, gb) = make_float2(acc[t][r][0], PAIR ? 0.f : acc[t][r][PAIR ? 0 : 1]);
}

} // namespace qbf
