// qbf_cnn.hip -- periodic 3x3 CNN surrogate of the collision operator: forward, backward and
// Adam kernels for gfx950 (no counterpart in the reference).
//
// Network: six 3x3 convolutions on the S x S velocity grid (S = N), periodic padding, GELU
// after the first five: 4 -> 64 -> 64 -> 64 -> 64 -> 64 -> 4 channels.  Activations are
// stored NHWC in bf16 (channels padded: 4 -> 32 at the input, 4 -> 32 at the output), weights
// as bf16 copies of fp32 master weights, and every convolution accumulates in fp32.
//
// The image side S is 64 or 128 (velocity grids N = 64, 128; a template parameter of the
// kernels).  Forward and backward-data convolutions are implicit GEMMs on the matrix cores
// (v_mfma_f32_16x16x32_bf16): each wave owns a 64-pixel segment of an image row times all output
// channels; K = 9 taps x input channels.
// MFMA operand layout (16 x 16 x 32): lane l holds A[l%16][8*(l/16) .. +7],
// B[8*(l/16) .. +7][l%16], and the result D[4*(l/16) + r][l%16], r < 4.
// The weight gradient (a reduction over all pixels) runs on the VALU in fp32.
// Options (qbf_cnn_configure): a dilation d_l per layer (tap offsets (ky-1, kx-1) * d_l), and
// per-cell normalisation: a_b = max |x[b][0]| (trace channel); the network sees x / a_b and the
// constant channel iscale * a_b, its output is scaled by oscale * a_b^2, and the training loss
// is taken on the normalised output.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/qbf.h"

namespace qbfcnn {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;
typedef __bf16 bf16;

constexpr int SEGP = 64;        // pixels of an image row segment (one wave of the convolutions)
constexpr int NL = 6;           // layers
constexpr int CH = 64;          // hidden channels
constexpr int CIO = 32;         // padded input / output channel count
constexpr int CREAL = 4;        // real input / output channels (Pauli components)

__host__ __device__ constexpr int cin_of(int l) { return l == 0 ? CIO : CH; }
__host__ __device__ constexpr int cout_of(int l) { return l == NL - 1 ? CIO : CH; }

// tanh-form GELU, z * 0.5 (1 + tanh u) with u = sqrt(2/pi) (z + 0.044715 z^3), evaluated through
// 0.5 (1 + tanh u) = 1 / (1 + exp(-2u)) with the hardware base-2 exponential and reciprocal (about
// 1e-7 relative error, far below the bf16 rounding of the activations; exp overflow gives 0, the
// z -> -inf limit)
__device__ __forceinline__ float half_one_plus_tanh(float u)
{
	const float e = __builtin_amdgcn_exp2f(-2.8853900817779268f * u);   // exp(-2u) = 2^(-2 log2(e) u)
	return __builtin_amdgcn_rcpf(1.0f + e);
}
__device__ __forceinline__ float gelu(float z)
{
	const float c = 0.7978845608028654f;   // sqrt(2/pi)
	const float u = c * (z + 0.044715f * z * z * z);
	return z * half_one_plus_tanh(u);
}
// derivative: s + z s (1 - s) 2 u', s = 0.5 (1 + tanh u), u' = sqrt(2/pi) (1 + 3 * 0.044715 z^2)
__device__ __forceinline__ float gelu_grad(float z)
{
	const float c = 0.7978845608028654f;
	const float u = c * (z + 0.044715f * z * z * z);
	const float sg = half_one_plus_tanh(u);
	return sg + 2.0f * z * sg * (1.0f - sg) * c * (1.0f + 3.0f * 0.044715f * z * z);
}

// a[b] = max |x[b][0][.]|  (one workgroup of 256 per cell)
template <int S>
__global__ __launch_bounds__(256) void k_cell_amp(const float* __restrict__ x, float* __restrict__ amp)
{
	__shared__ float sh[4];
	const float* xb = x + (size_t)blockIdx.x * CREAL * S * S;
	float m = 0.f;
	for (int p = threadIdx.x; p < S * S; p += 256) m = fmaxf(m, fabsf(xb[p]));
	for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
	if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
	__syncthreads();
	if (threadIdx.x == 0) amp[blockIdx.x] = fmaxf(fmaxf(sh[0], sh[1]), fmaxf(sh[2], sh[3]));
}

// planar fp32 [B][4][S][S] -> NHWC bf16 [B][S][S][CIO]: channels 0..3 = scale * x (or
// scale * x / a[b] with amp), channel 4 = scale * a[b] with amp, the rest zero
template <int S>
__global__ void k_to_nhwc(const float* __restrict__ x, bf16* __restrict__ out, int B, float scale,
	const float* __restrict__ amp)
{
	const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	const long long n = (long long)B * S * S;
	if (i >= n) return;
	const int b = (int)(i / (S * S)), p = (int)(i % (S * S));
	const float f = amp ? scale / fmaxf(amp[b], 1e-30f) : scale;
	bf16* o = out + i * CIO;
	for (int c = 0; c < CIO; c++) {
		float v = 0.0f;
		if (c < CREAL) v = x[((size_t)b * CREAL + c) * (S * S) + p] * f;
		else if (c == CREAL && amp) v = scale * amp[b];
		o[c] = (bf16)v;
	}
}

// Implicit-GEMM 3x3 periodic convolution on the matrix cores.
//   out[b][y][x][n] = sum_t sum_k in[b][y + DIR*d*(ky-1)][x + DIR*d*(kx-1)][k] * Wt[t][n][k]
// DIR = +1: forward (Wt = W[t][cout][cin]); DIR = -1: backward data (Wt = W^T[t][cin][cout]).
// MODE 0: out = bias + value, store pre-activation z (fp32) and GELU(z) (bf16 NHWC)
// MODE 1: out = bias + value, store fp32 planar [B][4][S][S] (first 4 channels), times
//         oscale (and a[b]^2 when amp is given)
// MODE 2: backward data: g = value * gelu'(zprev) -> bf16 NHWC (gradient of the previous z)
// A workgroup of CROWS waves computes CROWS consecutive row segments (64 pixels each: S / 64
// per image row); the layer's weights are staged once in LDS (row pitch KC + 4 elements: the 16
// lanes of a 16-byte read cover all 64 banks but one pair).  Each wave holds a 64-pixel x NC
// accumulator (4 x NC/16 MFMA tiles): per tap and 32-channel block it loads 4 A fragments
// and NC/16 B fragments from LDS for 4 * NC/16 MFMAs.
// ROWS (S = 64, forward, dilation 1): the workgroup's CROWS rows are consecutive rows of one image,
// so the CROWS + 2 input rows their taps read are staged in LDS once (each input row read once from
// global memory instead of nine times) and the A fragments come from LDS too.  Row r, pixel px,
// 16-byte channel chunk c sits at chunk slot (r S + px) CPP + (c ^ ((px / PPB) % CPP)) (CPP chunks
// per pixel, PPB pixels per 256 bytes): the 16 lanes of a fragment read, 16 consecutive pixels,
// then fall into 16 different 16-byte slots of a 256-byte bank line.  Otherwise the A fragments
// are loaded from global memory.
constexpr int CROWS = 8;
template <int S, int KC, int NC, int DIR, int MODE, bool ROWS = false>
__global__ __launch_bounds__(64 * CROWS) __attribute__((amdgpu_waves_per_eu(4))) void k_conv_mfma(const bf16* __restrict__ in, const bf16* __restrict__ Wt,
	const float* __restrict__ bias, float* __restrict__ zout, bf16* __restrict__ aout, float* __restrict__ yout,
	const float* __restrict__ zprev, float oscale, int dil, const float* __restrict__ amp)
{
	constexpr int NT = NC / 16;       // 16-wide output-channel tiles
	constexpr int MT = SEGP / 16;     // 16-pixel tiles of a segment
	constexpr int SEG = S / SEGP;     // segments per image row
	constexpr int KP = KC + 4;        // LDS pitch of a weight row (elements)
	constexpr int CPP = KC / 8;       // 16-byte channel chunks per pixel
	constexpr int PPB = 16 / CPP;     // pixels per 256 bytes
	static_assert(!ROWS || (S == SEGP && DIR == 1 && S % CROWS == 0), "row staging: S = 64, forward");
	__shared__ __attribute__((aligned(16))) bf16 wl[9 * NC * KP];
	__shared__ __attribute__((aligned(16))) bf16 xl[ROWS ? (CROWS + 2) * S * KC : 8];
	const int tile = blockIdx.x * CROWS + (threadIdx.x >> 6);
	const int b = tile / (S * SEG);
	const int y = (tile / SEG) % S;
	const int x0 = (tile % SEG) * SEGP;
	const int lane = threadIdx.x & 63;
	const int i16 = lane & 15, kq = lane >> 4;
	for (int v = threadIdx.x; v < 9 * NC * (KC / 8); v += 64 * CROWS) {
		const int r = v / (KC / 8), c8 = v % (KC / 8);
		const bf16x8 w = *(const bf16x8*)(Wt + (size_t)r * KC + c8 * 8);
		bf16* d = wl + r * KP + c8 * 8;              // 8-byte aligned (KP = KC + 4)
		*(uint2*)d = *(const uint2*)&w;
		*(uint2*)(d + 4) = *((const uint2*)&w + 1);
	}
	if constexpr (ROWS) {
		// input rows y0 - 1 .. y0 + CROWS of image b (periodic), y0 = the workgroup's first row
		const int y0 = (blockIdx.x * CROWS) % S;
		for (int q = threadIdx.x; q < (CROWS + 2) * S * CPP; q += 64 * CROWS) {
			const int r = q / (S * CPP), px = (q / CPP) % S, c = q % CPP;
			const int ys = (y0 - 1 + r + S) & (S - 1);
			const bf16x8 v = *(const bf16x8*)(in + (((size_t)b * S + ys) * S + px) * KC + c * 8);
			*(bf16x8*)(xl + ((r * S + px) * CPP + (c ^ ((px / PPB) % CPP))) * 8) = v;
		}
	}
	__syncthreads();
	f32x4 acc[MT][NT];
#pragma unroll
	for (int m = 0; m < MT; m++)
#pragma unroll
		for (int t = 0; t < NT; t++) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
	// K loop over the 9 taps x KC/32 channel blocks, the input fragments of the next iteration
	// loaded while the MFMAs of this one run (two iterations per loop body: the fragments ping-pong
	// between two register sets)
	constexpr int KB = KC / 32, NIT = 9 * KB;
	auto load_a = [&](int it, bf16x8* a) {
		const int tap = it / KB, kb = it % KB;
		if constexpr (ROWS) {
			// LDS row of the tap: the wave's row (y - y0) + tap / 3 (row 0 holds y0 - 1)
			const int r = (threadIdx.x >> 6) + tap / 3;
			const int c = kb * 4 + kq;
#pragma unroll
			for (int m = 0; m < MT; m++) {
				const int px = (m * 16 + i16 + tap % 3 - 1 + S) & (S - 1);
				a[m] = *(const bf16x8*)(xl + ((r * S + px) * CPP + (c ^ ((px / PPB) % CPP))) * 8);
			}
		} else {
			const int ys = (y + DIR * dil * (tap / 3 - 1) + 2 * S) & (S - 1);
			const int sh = DIR * dil * (tap % 3 - 1) + 2 * S;
			const bf16* rowp = in + ((size_t)b * S + ys) * S * KC + kb * 32 + kq * 8;
#pragma unroll
			for (int m = 0; m < MT; m++)
				a[m] = *(const bf16x8*)(rowp + (size_t)((x0 + m * 16 + i16 + sh) & (S - 1)) * KC);
		}
	};
	bf16x8 a[MT];
	load_a(0, a);
#pragma unroll 2
	for (int it = 0; it < NIT; it++) {
		const int tap = it / KB, kb = it % KB;
		const bf16* wt = wl + tap * NC * KP;
		bf16x8 w[NT];
#pragma unroll
		for (int t = 0; t < NT; t++) {
			const bf16* q = wt + (t * 16 + i16) * KP + kb * 32 + kq * 8;
			uint2 lo = *(const uint2*)q, hi = *(const uint2*)(q + 4);
			uint4 u = {lo.x, lo.y, hi.x, hi.y};
			w[t] = *(const bf16x8*)&u;
		}
		bf16x8 an[MT];
		load_a(it + 1 < NIT ? it + 1 : it, an);
#pragma unroll
		for (int m = 0; m < MT; m++)
#pragma unroll
			for (int t = 0; t < NT; t++) acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], w[t], acc[m][t], 0, 0, 0);
#pragma unroll
		for (int m = 0; m < MT; m++) a[m] = an[m];
	}
	// lane holds D[4*kq + r][i16] of tile (m, t): pixel m*16 + 4*kq + r, channel t*16 + i16
	float osc = oscale;
	if (MODE == 1 && amp) osc *= amp[b] * amp[b];
	if constexpr (MODE == 1) {
		// the CREAL output channels of the segment go through LDS (the weights are no longer
		// needed) so that the NCHW float rows are written whole: [c][64] per wave, 16 bytes a lane
		__syncthreads();
		float* ys = (float*)wl + (threadIdx.x >> 6) * (CREAL * SEGP);
		if (i16 < CREAL)
#pragma unroll
			for (int m = 0; m < MT; m++)
#pragma unroll
				for (int r = 0; r < 4; r++) ys[i16 * SEGP + m * 16 + kq * 4 + r] = osc * (acc[m][0][r] + bias[i16]);
		__syncthreads();
		static_assert(CREAL * SEGP == 4 * 64, "one float4 per lane");
		const int c = lane / (SEGP / 4), x4 = (lane % (SEGP / 4)) * 4;
		*(float4*)(yout + ((size_t)b * CREAL + c) * (S * S) + (size_t)y * S + x0 + x4) = *(const float4*)(ys + c * SEGP + x4);
		return;
	}
#pragma unroll
	for (int m = 0; m < MT; m++)
#pragma unroll
		for (int t = 0; t < NT; t++) {
			const int n = t * 16 + i16;
#pragma unroll
			for (int r = 0; r < 4; r++) {
				const int xo = x0 + m * 16 + kq * 4 + r;
				const size_t pix = ((size_t)b * S + y) * S + xo;
				const float v = acc[m][t][r];
				if (MODE == 0) {
					const float z = v + bias[n];
					if (zout) zout[pix * NC + n] = z;
					aout[pix * NC + n] = (bf16)gelu(z);
				} else if (MODE == 1) {
					if (n < CREAL) yout[((size_t)b * CREAL + n) * (S * S) + (size_t)y * S + xo] = osc * (v + bias[n]);
				} else {
					aout[pix * NC + n] = (bf16)(v * gelu_grad(zprev[pix * NC + n]));
				}
			}
		}
}

// Weight and bias gradients on the VALU, per workgroup:
//   part[blk][(t * CO + co) * CI + ci] = sum_p dz[p][co] * x[p + dil off_t][ci],
//   part[blk][9 CO CI + co] = sum_p dz[p][co]   (p over the workgroup's rows)
// summed over the workgroups in a fixed order by k_wgrad_reduce (no atomics: the gradients are the
// same bit for bit from run to run).  One workgroup per (image, block of ROWS rows); thread (co,
// 16-channel ci block) keeps 9 x 16 accumulators; the three rows of x that a row of dz meets and that
// row of dz are staged in LDS (periodic in x; the x rows with a pitch of CI + 4 floats, so that a
// thread reads its 16 channels as four aligned 16-byte words).
template <int S, int CI, int CO>
__global__ __launch_bounds__(256) void k_conv_wgrad(const bf16* __restrict__ x, const bf16* __restrict__ dz,
	float* __restrict__ part, int rows, int dil)
{
	constexpr int NCB = CI / 16;                 // ci blocks of 16
	constexpr int NTH = CO * NCB;                // active threads
	__shared__ __attribute__((aligned(16))) float xs[3][S][CI + 4];
	__shared__ float zs[S][CO + 1];
	const int b = blockIdx.x / (S / rows);
	const int y0 = (blockIdx.x % (S / rows)) * rows;
	const int tid = threadIdx.x;
	const int co = tid / NCB, cb = tid % NCB;
	float acc[9][16];
#pragma unroll
	for (int t = 0; t < 9; t++)
#pragma unroll
		for (int j = 0; j < 16; j++) acc[t][j] = 0.f;
	float bacc = 0.f;
	for (int y = y0; y < y0 + rows; y++) {
		__syncthreads();
		for (int e = tid; e < 3 * S * CI; e += blockDim.x) {
			const int r = e / (S * CI), rem = e % (S * CI);
			const int xx = rem / CI, c = rem % CI;
			const int ys = (y + dil * (r - 1) + 2 * S) & (S - 1);
			xs[r][xx][c] = (float)x[(((size_t)b * S + ys) * S + xx) * CI + c];
		}
		for (int e = tid; e < S * CO; e += blockDim.x) {
			const int xx = e / CO, c = e % CO;
			zs[xx][c] = (float)dz[(((size_t)b * S + y) * S + xx) * CO + c];
		}
		__syncthreads();
		if (tid < NTH) {
			for (int px = 0; px < S; px++) {
				const float g = zs[px][co];
				bacc += g;
#pragma unroll
				for (int kx = 0; kx < 3; kx++) {
					const int pxs = (px + dil * (kx - 1) + 2 * S) & (S - 1);
#pragma unroll
					for (int ky = 0; ky < 3; ky++)
#pragma unroll
						for (int j = 0; j < 16; j++) acc[ky * 3 + kx][j] += g * xs[ky][pxs][cb * 16 + j];
				}
			}
		}
	}
	if (tid < NTH) {
		float* pb = part + (size_t)blockIdx.x * (9 * CO * CI + CO);
#pragma unroll
		for (int t = 0; t < 9; t++)
#pragma unroll
			for (int j = 0; j < 16; j++) pb[((size_t)t * CO + co) * CI + cb * 16 + j] = acc[t][j];
		if (cb == 0) pb[9 * CO * CI + co] = bacc;
	}
}

// dW[i] = sum_blk part[blk][i] (i < nw), db[i - nw] likewise (nw <= i < nw + nb), blocks in order
__global__ void k_wgrad_reduce(const float* __restrict__ part, int nblk, int nw, int nb, float* __restrict__ dW,
	float* __restrict__ db)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	const int stride = nw + nb;
	if (i >= stride) return;
	float s = 0.f;
	for (int b = 0; b < nblk; b++) s += part[(size_t)b * stride + i];
	if (i < nw) dW[i] = s;
	else db[i - nw] = s;
}

// loss = mean over elements of d^2 (per-workgroup partials into lpart, summed by k_loss_reduce),
// gradient with respect to the network output -> bf16 NHWC (CIO channels).  d = pred - target, or
// (pred - target) / (oscale a[b]^2) with amp.
template <int S>
__global__ __launch_bounds__(256) void k_loss_grad(const float* __restrict__ pred, const float* __restrict__ target,
	bf16* __restrict__ g, float* __restrict__ lpart, int B, float inv_n, float oscale, const float* __restrict__ amp)
{
	const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	const long long n = (long long)B * S * S;
	float part = 0.f;
	if (i < n) {
		const int b = (int)(i / (S * S)), p = (int)(i % (S * S));
		const float sc = amp ? oscale * amp[b] * amp[b] : 1.0f;
		bf16* o = g + i * CIO;
		for (int c = 0; c < CIO; c++) {
			float gv = 0.f;
			if (c < CREAL) {
				const size_t e = ((size_t)b * CREAL + c) * (S * S) + p;
				const float d = amp ? (pred[e] - target[e]) / sc : pred[e] - target[e];
				part += d * d;
				gv = amp ? 2.0f * d * inv_n : 2.0f * d * inv_n * oscale;   // d loss / d (network output)
			}
			o[c] = (bf16)gv;
		}
	}
	for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o);
	__shared__ float wsum[4];
	if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = part;
	__syncthreads();
	if (threadIdx.x == 0) lpart[blockIdx.x] = ((wsum[0] + wsum[1]) + (wsum[2] + wsum[3])) * inv_n;
}

// loss = sum of the per-block partials, in order (one workgroup of 256)
__global__ __launch_bounds__(256) void k_loss_reduce(const float* __restrict__ lpart, int n, float* __restrict__ loss)
{
	__shared__ float sh[256];
	float s = 0.f;
	for (int i = threadIdx.x; i < n; i += 256) s += lpart[i];
	sh[threadIdx.x] = s;
	__syncthreads();
	for (int w = 128; w > 0; w >>= 1) {
		if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
		__syncthreads();
	}
	if (threadIdx.x == 0) *loss = sh[0];
}

// Adam on fp32 master weights, then bf16 copies W[t][co][ci] and W^T[t][ci][co].
__global__ void k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v, const float* __restrict__ g,
	bf16* __restrict__ wb, bf16* __restrict__ wbt, int co_n, int ci_n, int co_real, int ci_real, int n,
	float lr, float b1, float b2, float eps, float bc1, float bc2, int is_bias)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	int co, ci, t;
	if (is_bias) { co = i; ci = 0; t = 0; }
	else { t = i / (co_n * ci_n); co = (i / ci_n) % co_n; ci = i % ci_n; }
	const bool live = co < co_real && (is_bias || ci < ci_real);
	float wv = w[i];
	if (live) {
		const float gi = g[i];
		const float mi = b1 * m[i] + (1 - b1) * gi;
		const float vi = b2 * v[i] + (1 - b2) * gi * gi;
		m[i] = mi;
		v[i] = vi;
		wv -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
	} else {
		wv = 0.f;
	}
	w[i] = wv;
	if (!is_bias) {
		wb[i] = (bf16)wv;
		wbt[((size_t)t * ci_n + ci) * co_n + co] = (bf16)wv;
	}
}

// bf16 copies of fp32 weights (after an upload)
__global__ void k_wcopy(const float* __restrict__ w, bf16* __restrict__ wb, bf16* __restrict__ wbt, int co_n, int ci_n, int n)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const int t = i / (co_n * ci_n), co = (i / ci_n) % co_n, ci = i % ci_n;
	wb[i] = (bf16)w[i];
	wbt[((size_t)t * ci_n + ci) * co_n + co] = (bf16)w[i];
}

// surrogate midpoint step, elementwise over the [ncell][4][S][S] state; products and sums are
// rounded separately (no fused multiply-add), so the step equals its composition from single
// network evaluations bit for bit
__global__ void k_to_f32(const double* __restrict__ W, float* __restrict__ x, long long n)
{
	const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) x[i] = (float)W[i];
}
// C = Cv (+ D): the network's operator, plus the offset field D when one is set
__device__ __forceinline__ double op_value(const float* __restrict__ Cv, const float* __restrict__ D, long long i)
{
#pragma clang fp contract(off)
	return D ? (double)Cv[i] + (double)D[i] : (double)Cv[i];
}
// x = (float)(W + a C)
__global__ void k_half_f32(const double* __restrict__ W, const float* __restrict__ Cv, const float* __restrict__ D,
	double a, float* __restrict__ x, long long n)
{
#pragma clang fp contract(off)
	const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) x[i] = (float)(W[i] + a * op_value(Cv, D, i));
}
// out = W + a C
__global__ void k_update_f64(const double* __restrict__ W, const float* __restrict__ Cv, const float* __restrict__ D,
	double a, double* __restrict__ out, long long n)
{
#pragma clang fp contract(off)
	const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) out[i] = W[i] + a * op_value(Cv, D, i);
}

// gather cells idx[j] of W [*][CELL] into out [nsel][CELL] (CELL = 4 S S values)
template <typename V>
__global__ void k_gather_cells(const V* __restrict__ W, const int* __restrict__ idx, V* __restrict__ out, int nsel,
	int CELL)
{
	const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= (long long)nsel * CELL) return;
	const int j = (int)(i / CELL), e = (int)(i % CELL);
	out[i] = W[(size_t)idx[j] * CELL + e];
}

// res[0] = sum |a (+ D) - b|, res[1] = sum |b|, res[2] = sum |w|  (a and D fp32, b and w fp64), in two
// launches with a fixed summation order (the same sums bit for bit from run to run): RES_WG workgroups
// of 256 write partial sums part[wg][3] over a grid-stride range, then one workgroup adds them in order.
constexpr int RES_WG = 128;
__global__ __launch_bounds__(256) void k_l1_partial(const float* __restrict__ a, const float* __restrict__ D,
	const double* __restrict__ b, const double* __restrict__ w, long long n, double* __restrict__ part)
{
	__shared__ double sh[3][4];
	double d = 0, m = 0, s = 0;
	for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
		d += fabs(op_value(a, D, i) - b[i]);
		m += fabs(b[i]);
		s += fabs(w[i]);
	}
	for (int o = 32; o > 0; o >>= 1) {
		d += __shfl_xor(d, o);
		m += __shfl_xor(m, o);
		s += __shfl_xor(s, o);
	}
	if ((threadIdx.x & 63) == 0) {
		sh[0][threadIdx.x >> 6] = d;
		sh[1][threadIdx.x >> 6] = m;
		sh[2][threadIdx.x >> 6] = s;
	}
	__syncthreads();
	if (threadIdx.x < 3) {
		const int k = threadIdx.x;
		part[blockIdx.x * 3 + k] = ((sh[k][0] + sh[k][1]) + sh[k][2]) + sh[k][3];
	}
}

__global__ __launch_bounds__(64) void k_l1_final(const double* __restrict__ part, int nwg, double* __restrict__ res)
{
	if (threadIdx.x < 3) {
		double t = 0;
		for (int g = 0; g < nwg; g++) t += part[g * 3 + threadIdx.x];
		res[threadIdx.x] = t;
	}
}

} // namespace qbfcnn

using namespace qbfcnn;

static thread_local std::string g_cnn_err;
static int cnn_err(int c, const char* fmt, ...)
{
	char buf[512];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof(buf), fmt, ap);
	va_end(ap);
	g_cnn_err = buf;
	return c;
}
#define HCHK(x)                                                              \
	do {                                                                     \
		hipError_t e_ = (x);                                                 \
		if (e_ != hipSuccess) return cnn_err(-2, "%s", hipGetErrorString(e_));  \
	} while (0)

struct qbf_cnn {
	int S = 64;                      // image side (the velocity grid N: 64 or 128)
	int maxB = 0;
	// training (qbf_cnn_gradients) works on batches of at most min(maxB, TRAIN_MAX) images: the
	// pre-activations and gradients are held for those only (inference batches may be larger)
	static constexpr int TRAIN_MAX = 256;
	int maxTrain = 0;
	hipStream_t own = nullptr;       // created by qbf_cnn_create, destroyed by qbf_cnn_destroy
	hipStream_t stream = nullptr;    // where work is launched (own, or the caller's)
	// per layer: fp32 master weights / grads / Adam moments, bf16 copies
	float *w[NL] = {}, *gw[NL] = {}, *mw[NL] = {}, *vw[NL] = {};
	float *bias[NL] = {}, *gb[NL] = {}, *mb[NL] = {}, *vb[NL] = {};
	bf16 *wb[NL] = {}, *wbt[NL] = {};
	// activations: x[l] = input of layer l (bf16 NHWC), z[l] = pre-activation (fp32 NHWC), l < NL-1
	bf16* x[NL] = {};
	float* z[NL] = {};
	bf16* g[NL] = {};     // gradient of z[l] (bf16 NHWC), g[NL-1] = gradient of the output
	float* loss = nullptr;
	// deterministic reductions: per-workgroup partials of the weight gradients of one layer and of the loss
	static constexpr int WG_ROWS = 16;    // image rows per workgroup of k_conv_wgrad
	float* gpart = nullptr;
	float* lpart = nullptr;
	int dil[NL] = {1, 1, 1, 1, 1, 1};
	int norm = 0;                         // per-cell normalisation
	float* amp = nullptr;                 // per-cell amplitude [maxB]
	float *sx = nullptr, *sc = nullptr;   // surrogate step: network input / output, fp32 [maxB][4][S][S]
	double* sw = nullptr;                 // surrogate inhom step: state after the first half transport
	double* sw2 = nullptr;                // ... and after the collision step
	static constexpr int RMAX = 64;       // residual check: at most RMAX sampled cells
	int* sidx = nullptr;                  // sampled cell indices (device)
	double *swf = nullptr, *swc = nullptr; // Fourier state / exact collision of the sample (complex)
	// offset field D (fp32 [ncell][4][S][S], the caller's): the surrogate operator is net + D
	const float* off = nullptr;
	int offCells = 0;                     // the cell count D was given for
	float* soff = nullptr;                // D on the residual sample
	double* sres = nullptr;               // three sums of the residual, then their per-workgroup partials
	int step = 0;
	std::vector<void*> allocs;
	float* alloc_f(size_t n)
	{
		void* p = nullptr;
		if (hipMalloc(&p, n * sizeof(float)) != hipSuccess) return nullptr;
		allocs.push_back(p);
		if (hipMemset(p, 0, n * sizeof(float)) != hipSuccess) return nullptr;
		return (float*)p;
	}
	bf16* alloc_b(size_t n)
	{
		void* p = nullptr;
		if (hipMalloc(&p, n * sizeof(bf16)) != hipSuccess) return nullptr;
		allocs.push_back(p);
		if (hipMemset(p, 0, n * sizeof(bf16)) != hipSuccess) return nullptr;
		return (bf16*)p;
	}
	~qbf_cnn()
	{
		for (void* p : allocs) hipFree(p);
	}
};

static size_t wsize(int l) { return (size_t)9 * cout_of(l) * cin_of(l); }

extern "C" {

const char* qbf_cnn_last_error(void) { return g_cnn_err.c_str(); }

int qbf_cnn_create(int N, int maxBatch, qbf_cnn** out)
{
	if (!out || (N != 64 && N != 128) || maxBatch < 1)
		return cnn_err(-1, "qbf_cnn_create: N must be 64 or 128 and maxBatch >= 1");
	qbf_cnn* c = new qbf_cnn();
	const int S = N;
	c->S = N;
	c->maxB = maxBatch;
	c->maxTrain = maxBatch < qbf_cnn::TRAIN_MAX ? maxBatch : qbf_cnn::TRAIN_MAX;
	const size_t px = (size_t)maxBatch * S * S, pxt = (size_t)c->maxTrain * S * S;
	for (int l = 0; l < NL; l++) {
		const size_t nw = wsize(l), nb = cout_of(l);
		c->w[l] = c->alloc_f(nw); c->gw[l] = c->alloc_f(nw); c->mw[l] = c->alloc_f(nw); c->vw[l] = c->alloc_f(nw);
		c->bias[l] = c->alloc_f(nb); c->gb[l] = c->alloc_f(nb); c->mb[l] = c->alloc_f(nb); c->vb[l] = c->alloc_f(nb);
		c->wb[l] = c->alloc_b(nw); c->wbt[l] = c->alloc_b(nw);
		c->x[l] = c->alloc_b(px * cin_of(l));
		c->g[l] = c->alloc_b(pxt * cout_of(l));
		if (l < NL - 1) c->z[l] = c->alloc_f(pxt * cout_of(l));
		if (!c->w[l] || !c->gw[l] || !c->mw[l] || !c->vw[l] || !c->bias[l] || !c->gb[l] || !c->mb[l] || !c->vb[l] ||
			!c->wb[l] || !c->wbt[l] || !c->x[l] || !c->g[l] || (l < NL - 1 && !c->z[l])) {
			delete c;
			return cnn_err(-3, "qbf_cnn_create: out of device memory");
		}
	}
	c->loss = c->alloc_f(1);
	c->gpart = c->alloc_f((size_t)c->maxTrain * (S / qbf_cnn::WG_ROWS) * (9 * CH * CH + CH));
	c->lpart = c->alloc_f(pxt / 256 + 1);
	c->amp = c->alloc_f(maxBatch);
	c->sx = c->alloc_f(px * CREAL);
	c->sc = c->alloc_f(px * CREAL);
	c->soff = c->alloc_f((size_t)qbf_cnn::RMAX * CREAL * S * S);
	{
		void *p1 = nullptr, *p2 = nullptr;
		if (hipMalloc(&p1, px * CREAL * sizeof(double)) == hipSuccess) c->allocs.push_back(p1);
		else p1 = nullptr;
		if (hipMalloc(&p2, px * CREAL * sizeof(double)) == hipSuccess) c->allocs.push_back(p2);
		else p2 = nullptr;
		c->sw = (double*)p1;
		c->sw2 = (double*)p2;
	}
	{
		void *p1 = nullptr, *p2 = nullptr, *p3 = nullptr, *p4 = nullptr;
		const size_t cplx = (size_t)qbf_cnn::RMAX * CREAL * S * S * 2 * sizeof(double);
		if (hipMalloc(&p1, qbf_cnn::RMAX * sizeof(int)) == hipSuccess) c->allocs.push_back(p1); else p1 = nullptr;
		if (hipMalloc(&p2, cplx) == hipSuccess) c->allocs.push_back(p2); else p2 = nullptr;
		if (hipMalloc(&p3, cplx) == hipSuccess) c->allocs.push_back(p3); else p3 = nullptr;
		// the three sums, then RES_WG partial sums of each
		if (hipMalloc(&p4, 3 * (1 + RES_WG) * sizeof(double)) == hipSuccess) c->allocs.push_back(p4); else p4 = nullptr;
		c->sidx = (int*)p1;
		c->swf = (double*)p2;
		c->swc = (double*)p3;
		c->sres = (double*)p4;
	}
	if (!c->loss || !c->gpart || !c->lpart || !c->amp || !c->sx || !c->sc || !c->soff || !c->sw || !c->sw2 || !c->sidx || !c->swf || !c->swc || !c->sres) {
		delete c;
		return cnn_err(-3, "qbf_cnn_create: out of device memory");
	}
	if (hipStreamCreate(&c->own) != hipSuccess) {
		delete c;
		return cnn_err(-2, "hipStreamCreate failed");
	}
	c->stream = c->own;
	*out = c;
	return 0;
}

void qbf_cnn_destroy(qbf_cnn* c)
{
	if (!c) return;
	hipStreamSynchronize(c->stream);
	hipStreamDestroy(c->own);
	delete c;
}

int qbf_cnn_set_stream(qbf_cnn* c, void* s)
{
	if (!c) return cnn_err(-1, "bad argument");
	c->stream = s ? (hipStream_t)s : c->own;
	return 0;
}

// dilation per layer (1 <= d <= 16; NULL keeps the current ones) and per-cell normalisation (0/1)
int qbf_cnn_configure(qbf_cnn* c, const int* dil, int norm)
{
	if (!c || (norm != 0 && norm != 1)) return cnn_err(-1, "qbf_cnn_configure: bad argument");
	if (dil)
		for (int l = 0; l < NL; l++)
			if (dil[l] < 1 || dil[l] > 16) return cnn_err(-1, "qbf_cnn_configure: dilation must be in [1, 16]");
	if (dil)
		for (int l = 0; l < NL; l++) c->dil[l] = dil[l];
	c->norm = norm;
	return 0;
}

// parameter layout of the host buffer: for each layer, weights [9][cout_pad][cin_pad] then bias [cout_pad]
size_t qbf_cnn_param_count(void)
{
	size_t n = 0;
	for (int l = 0; l < NL; l++) n += wsize(l) + cout_of(l);
	return n;
}

int qbf_cnn_set_params(qbf_cnn* c, const float* host)
{
	if (!c || !host) return cnn_err(-1, "bad argument");
	size_t o = 0;
	for (int l = 0; l < NL; l++) {
		HCHK(hipMemcpy(c->w[l], host + o, wsize(l) * sizeof(float), hipMemcpyHostToDevice));
		o += wsize(l);
		HCHK(hipMemcpy(c->bias[l], host + o, cout_of(l) * sizeof(float), hipMemcpyHostToDevice));
		o += cout_of(l);
		const int n = (int)wsize(l);
		hipLaunchKernelGGL(k_wcopy, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->w[l], c->wb[l], c->wbt[l],
			cout_of(l), cin_of(l), n);
		HCHK(hipGetLastError());
		HCHK(hipMemsetAsync(c->mw[l], 0, wsize(l) * sizeof(float), c->stream));
		HCHK(hipMemsetAsync(c->vw[l], 0, wsize(l) * sizeof(float), c->stream));
		HCHK(hipMemsetAsync(c->mb[l], 0, cout_of(l) * sizeof(float), c->stream));
		HCHK(hipMemsetAsync(c->vb[l], 0, cout_of(l) * sizeof(float), c->stream));
	}
	c->step = 0;
	HCHK(hipStreamSynchronize(c->stream));
	return 0;
}

int qbf_cnn_get_params(qbf_cnn* c, float* host, int grads)
{
	if (!c || !host) return cnn_err(-1, "bad argument");
	HCHK(hipStreamSynchronize(c->stream));
	size_t o = 0;
	for (int l = 0; l < NL; l++) {
		HCHK(hipMemcpy(host + o, grads ? c->gw[l] : c->w[l], wsize(l) * sizeof(float), hipMemcpyDeviceToHost));
		o += wsize(l);
		HCHK(hipMemcpy(host + o, grads ? c->gb[l] : c->bias[l], cout_of(l) * sizeof(float), hipMemcpyDeviceToHost));
		o += cout_of(l);
	}
	return 0;
}

} // extern "C"

// convolution grid: CROWS row segments of 64 pixels per workgroup
template <int S>
static dim3 conv_grid(int B) { return dim3((unsigned)(B * S * (S / SEGP) / CROWS)); }

template <int S>
static int run_forward_t(qbf_cnn* c, const float* dIn, int B, float iscale, float* dOut, float oscale, bool train)
{
	const long long npx = (long long)B * S * S;
	const float* amp = c->norm ? c->amp : nullptr;
	if (c->norm) hipLaunchKernelGGL(k_cell_amp<S>, dim3(B), dim3(256), 0, c->stream, dIn, c->amp);
	hipLaunchKernelGGL(k_to_nhwc<S>, dim3((unsigned)((npx + 255) / 256)), dim3(256), 0, c->stream, dIn, c->x[0], B,
		iscale, amp);
	const dim3 grid = conv_grid<S>(B), blk(64 * CROWS);
	// the input rows staged in LDS where the taps allow it (S = 64, dilation 1)
	constexpr bool R = S == SEGP;
	if (R && c->dil[0] == 1)
		hipLaunchKernelGGL((k_conv_mfma<S, CIO, CH, 1, 0, R>), grid, blk, 0, c->stream, c->x[0], c->wb[0], c->bias[0],
			train ? c->z[0] : nullptr, c->x[1], nullptr, nullptr, 1.0f, 1, nullptr);
	else
		hipLaunchKernelGGL((k_conv_mfma<S, CIO, CH, 1, 0>), grid, blk, 0, c->stream, c->x[0], c->wb[0], c->bias[0],
			train ? c->z[0] : nullptr, c->x[1], nullptr, nullptr, 1.0f, c->dil[0], nullptr);
	for (int l = 1; l < NL - 1; l++) {
		if (R && c->dil[l] == 1)
			hipLaunchKernelGGL((k_conv_mfma<S, CH, CH, 1, 0, R>), grid, blk, 0, c->stream, c->x[l], c->wb[l], c->bias[l],
				train ? c->z[l] : nullptr, c->x[l + 1], nullptr, nullptr, 1.0f, 1, nullptr);
		else
			hipLaunchKernelGGL((k_conv_mfma<S, CH, CH, 1, 0>), grid, blk, 0, c->stream, c->x[l], c->wb[l], c->bias[l],
				train ? c->z[l] : nullptr, c->x[l + 1], nullptr, nullptr, 1.0f, c->dil[l], nullptr);
	}
	if (R && c->dil[NL - 1] == 1)
		hipLaunchKernelGGL((k_conv_mfma<S, CH, CIO, 1, 1, R>), grid, blk, 0, c->stream, c->x[NL - 1], c->wb[NL - 1],
			c->bias[NL - 1], nullptr, nullptr, dOut, nullptr, oscale, 1, amp);
	else
		hipLaunchKernelGGL((k_conv_mfma<S, CH, CIO, 1, 1>), grid, blk, 0, c->stream, c->x[NL - 1], c->wb[NL - 1],
			c->bias[NL - 1], nullptr, nullptr, dOut, nullptr, oscale, c->dil[NL - 1], amp);
	HCHK(hipGetLastError());
	return 0;
}

extern "C" {

static int run_forward(qbf_cnn* c, const float* dIn, int B, float iscale, float* dOut, float oscale, bool train)
{
	return c->S == 128 ? run_forward_t<128>(c, dIn, B, iscale, dOut, oscale, train)
	                   : run_forward_t<64>(c, dIn, B, iscale, dOut, oscale, train);
}

// out = oscale * net(iscale * in), planar fp32 [B][4][N][N] device arrays
int qbf_cnn_forward(qbf_cnn* c, const float* dIn, int B, float iscale, float* dOut, float oscale)
{
	if (!c || !dIn || !dOut || B < 1 || B > c->maxB) return cnn_err(-1, "qbf_cnn_forward: bad argument");
	return run_forward(c, dIn, B, iscale, dOut, oscale, false);
}

// gradients of loss = mean (out - target)^2 into the gradient buffers (no update); loss -> *hostLoss
} // extern "C"

template <int S>
static int gradients_t(qbf_cnn* c, const float* dIn, const float* dTarget, int B, float iscale, float oscale,
	float* dOut, float* hostLoss)
{
	int r = run_forward_t<S>(c, dIn, B, iscale, dOut, oscale, true);
	if (r) return r;
	const long long npx = (long long)B * S * S;
	const float inv_n = 1.0f / (float)(npx * CREAL);
	const int nlb = (int)((npx + 255) / 256);
	hipLaunchKernelGGL(k_loss_grad<S>, dim3((unsigned)nlb), dim3(256), 0, c->stream, dOut, dTarget, c->g[NL - 1],
		c->lpart, B, inv_n, oscale, c->norm ? c->amp : nullptr);
	hipLaunchKernelGGL(k_loss_reduce, dim3(1), dim3(256), 0, c->stream, c->lpart, nlb, c->loss);
	const dim3 grid = conv_grid<S>(B), cblk(64 * CROWS), blk(256);
	const int rows = qbf_cnn::WG_ROWS;
	const int nwb = B * (S / rows);
	const dim3 wgrid(nwb);
	auto reduce = [&](int l) {
		const int nw = (int)wsize(l), nb = cout_of(l);
		hipLaunchKernelGGL(k_wgrad_reduce, dim3((unsigned)((nw + nb + 255) / 256)), dim3(256), 0, c->stream, c->gpart,
			nwb, nw, nb, c->gw[l], c->gb[l]);
	};
	// last layer: weight gradient, then data gradient into the last hidden pre-activation
	hipLaunchKernelGGL((k_conv_wgrad<S, CH, CIO>), wgrid, blk, 0, c->stream, c->x[NL - 1], c->g[NL - 1], c->gpart, rows,
		c->dil[NL - 1]);
	reduce(NL - 1);
	hipLaunchKernelGGL((k_conv_mfma<S, CIO, CH, -1, 2>), grid, cblk, 0, c->stream, c->g[NL - 1], c->wbt[NL - 1], nullptr,
		nullptr, c->g[NL - 2], nullptr, c->z[NL - 2], 1.0f, c->dil[NL - 1], nullptr);
	for (int l = NL - 2; l >= 1; l--) {
		hipLaunchKernelGGL((k_conv_wgrad<S, CH, CH>), wgrid, blk, 0, c->stream, c->x[l], c->g[l], c->gpart, rows,
			c->dil[l]);
		reduce(l);
		hipLaunchKernelGGL((k_conv_mfma<S, CH, CH, -1, 2>), grid, cblk, 0, c->stream, c->g[l], c->wbt[l], nullptr, nullptr,
			c->g[l - 1], nullptr, c->z[l - 1], 1.0f, c->dil[l], nullptr);
	}
	hipLaunchKernelGGL((k_conv_wgrad<S, CIO, CH>), wgrid, blk, 0, c->stream, c->x[0], c->g[0], c->gpart, rows, c->dil[0]);
	reduce(0);
	HCHK(hipGetLastError());
	if (hostLoss) {
		HCHK(hipMemcpyAsync(hostLoss, c->loss, sizeof(float), hipMemcpyDeviceToHost, c->stream));
		HCHK(hipStreamSynchronize(c->stream));
	}
	return 0;
}

extern "C" {

int qbf_cnn_gradients(qbf_cnn* c, const float* dIn, const float* dTarget, int B, float iscale, float oscale,
	float* dOut, float* hostLoss)
{
	if (!c || !dIn || !dTarget || !dOut || B < 1 || B > c->maxTrain)
		return cnn_err(-1, "qbf_cnn_gradients: bad argument (1 <= B <= min(maxBatch, 256))");
	return c->S == 128 ? gradients_t<128>(c, dIn, dTarget, B, iscale, oscale, dOut, hostLoss)
	                   : gradients_t<64>(c, dIn, dTarget, B, iscale, oscale, dOut, hostLoss);
}

// one Adam step with the gradients of qbf_cnn_gradients
int qbf_cnn_adam(qbf_cnn* c, float lr, float b1, float b2, float eps)
{
	if (!c) return cnn_err(-1, "null cnn");
	c->step++;
	const float bc1 = 1.0f - powf(b1, (float)c->step), bc2 = 1.0f - powf(b2, (float)c->step);
	for (int l = 0; l < NL; l++) {
		const int co = cout_of(l), ci = cin_of(l);
		const int co_r = l == NL - 1 ? CREAL : CH, ci_r = l == 0 ? (c->norm ? CREAL + 1 : CREAL) : CH;
		const int n = (int)wsize(l);
		hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->w[l], c->mw[l], c->vw[l], c->gw[l],
			c->wb[l], c->wbt[l], co, ci, co_r, ci_r, n, lr, b1, b2, eps, bc1, bc2, 0);
		hipLaunchKernelGGL(k_adam, dim3((co + 255) / 256), dim3(256), 0, c->stream, c->bias[l], c->mb[l], c->vb[l],
			c->gb[l], nullptr, nullptr, co, 1, co_r, 1, co, lr, b1, b2, eps, bc1, bc2, 1);
	}
	HCHK(hipGetLastError());
	return 0;
}

// Midpoint collision step with the network in place of the collision operator, on physical
// states [ncell][4][N][N] (double, device):
//   C(W) ~ oscale * net(iscale * W),  W_half = W + dt/2 C(W),  out = W + dt C(W_half),
// then the FP64 moment correction of out - W (mask as qbf_moment_correct; 0 = none).
// Runs on the engine's stream; ncell may exceed maxBatch (processed in chunks).
int qbf_surrogate_collision_step(qbf_cnn* c, qbf_engine* e, const double* dW, int ncell, double dt, float iscale,
	float oscale, double L, int mask, double* dOut)
{
	if (!c || !e || !dW || !dOut || ncell < 1) return cnn_err(-1, "qbf_surrogate_collision_step: bad argument");
	if (qbf_engine_N(e) != c->S) return cnn_err(-1, "qbf_surrogate_collision_step: engine and network grids differ");
	if (c->off && ncell != c->offCells)
		return cnn_err(-1, "qbf_surrogate_collision_step: %d cells, but the offset field holds %d", ncell, c->offCells);
	const int S = c->S;
	if (dW == dOut) return cnn_err(-1, "qbf_surrogate_collision_step: in-place update not supported");
	hipStream_t saved = c->stream;
	c->stream = (hipStream_t)qbf_get_stream(e);
	int r = 0;
	for (int c0 = 0; c0 < ncell && !r; c0 += c->maxB) {
		const int nb = ncell - c0 < c->maxB ? ncell - c0 : c->maxB;
		const long long n = (long long)nb * CREAL * S * S;
		const size_t off = (size_t)c0 * CREAL * S * S;
		const dim3 g((unsigned)((n + 255) / 256)), b(256);
		hipLaunchKernelGGL(k_to_f32, g, b, 0, c->stream, dW + off, c->sx, n);
		r = run_forward(c, c->sx, nb, iscale, c->sc, oscale, false);
		if (r) break;
		const float* D = c->off ? c->off + off : nullptr;
		hipLaunchKernelGGL(k_half_f32, g, b, 0, c->stream, dW + off, c->sc, D, 0.5 * dt, c->sx, n);
		r = run_forward(c, c->sx, nb, iscale, c->sc, oscale, false);
		if (r) break;
		hipLaunchKernelGGL(k_update_f64, g, b, 0, c->stream, dW + off, c->sc, D, dt, dOut + off, n);
		if (hipGetLastError() != hipSuccess) r = cnn_err(-2, "surrogate step launch failed");
	}
	if (!r && mask && qbf_moment_correct(e, dW, dOut, ncell, L, mask))
		r = cnn_err(-2, "%s", qbf_last_error());
	c->stream = saved;
	return r;
}

// Residual of the surrogate on sampled cells of physical states W: the exact collision
// operator C = Cc + Cd of the engine (any precision) against C_net = oscale * net(iscale * W),
// summed over the nsample cells hostIdx[0..nsample-1]:
//   sums[0] = sum |C_net - C|,  sums[1] = sum |C|,  sums[2] = sum |W|.
// sums[0] / sums[1] is the relative error of the operator; dt * sums[0] / sums[2] the relative
// error one surrogate step adds to the state.  Synchronous (the sums are read back to the host).
int qbf_surrogate_residual(qbf_cnn* c, qbf_engine* e, const double* dW, int ncell, const int* hostIdx, int nsample,
	float iscale, float oscale, double* sums)
{
	if (!c || !e || !dW || !hostIdx || !sums || nsample < 1 || nsample > qbf_cnn::RMAX || nsample > c->maxB)
		return cnn_err(-1, "qbf_surrogate_residual: bad argument (1 <= nsample <= 64 and <= maxBatch)");
	if (qbf_engine_N(e) != c->S) return cnn_err(-1, "qbf_surrogate_residual: engine and network grids differ");
	if (c->off && ncell != c->offCells)
		return cnn_err(-1, "qbf_surrogate_residual: %d cells, but the offset field holds %d", ncell, c->offCells);
	const int S = c->S;
	for (int j = 0; j < nsample; j++)
		if (hostIdx[j] < 0 || hostIdx[j] >= ncell) return cnn_err(-1, "qbf_surrogate_residual: cell index out of range");
	hipStream_t saved = c->stream;
	c->stream = (hipStream_t)qbf_get_stream(e);
	int r = 0;
	const long long n = (long long)nsample * CREAL * S * S;
	double h[3] = {0, 0, 0};
	if (hipMemcpyAsync(c->sidx, hostIdx, nsample * sizeof(int), hipMemcpyHostToDevice, c->stream) != hipSuccess)
		r = cnn_err(-2, "index upload failed");
	if (!r) {
		hipLaunchKernelGGL(k_gather_cells<double>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, dW, c->sidx,
			c->sw, nsample, CREAL * S * S);
		if (c->off)
			hipLaunchKernelGGL(k_gather_cells<float>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->off,
				c->sidx, c->soff, nsample, CREAL * S * S);
		if (qbf_phys_to_fourier(e, c->sw, CREAL * nsample, c->swf) ||
			qbf_collision(e, c->swf, nsample, QBF_C, c->swc, nullptr, nullptr, nullptr, 1.0) ||
			qbf_fourier_to_phys(e, c->swc, CREAL * nsample, c->sw2))
			r = cnn_err(-2, "%s", qbf_last_error());
	}
	if (!r) {
		hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->sw, c->sx, n);
		r = run_forward(c, c->sx, nsample, iscale, c->sc, oscale, false);
	}
	if (!r) {
		hipLaunchKernelGGL(k_l1_partial, dim3(RES_WG), dim3(256), 0, c->stream, c->sc, c->off ? c->soff : nullptr,
			c->sw2, c->sw, n, c->sres + 3);
		hipLaunchKernelGGL(k_l1_final, dim3(1), dim3(64), 0, c->stream, c->sres + 3, RES_WG, c->sres);
		if (hipMemcpyAsync(h, c->sres, sizeof(h), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
			hipStreamSynchronize(c->stream) != hipSuccess)
			r = cnn_err(-2, "residual read-back failed");
	}
	c->stream = saved;
	if (!r)
		for (int k = 0; k < 3; k++) sums[k] = h[k];
	return r;
}

// Offset field of the surrogate operator: C(W) ~ oscale * net(iscale * W) + D (D fp32, planar
// [ncell][4][N][N] device array, the same cells as the states passed to the surrogate calls; the
// pointer is kept until the next call; NULL: no offset).  The stepping and residual calls reject
// another cell count while D is set.
int qbf_surrogate_set_offset(qbf_cnn* c, const float* dD, int ncell)
{
	if (!c) return cnn_err(-1, "qbf_surrogate_set_offset: null network");
	if (dD && ncell < 1) return cnn_err(-1, "qbf_surrogate_set_offset: ncell must be >= 1");
	c->off = dD;
	c->offCells = dD ? ncell : 0;
	return 0;
}

// Strang step of the periodic inhomogeneous problem with the surrogate collision step:
// half transport (FP64, exact), surrogate midpoint collision step + moment correction, half
// transport.  ncell <= maxBatch (the intermediate states live in the network's buffers).
int qbf_surrogate_inhom_step_periodic(qbf_cnn* c, qbf_engine* e, const double* dW, int ncell, double L, double h,
	double dt, float iscale, float oscale, int mask, double* dOut)
{
	if (!c || !e || !dW || !dOut || ncell < 1 || ncell > c->maxB)
		return cnn_err(-1, "qbf_surrogate_inhom_step_periodic: bad argument (ncell must be in [1, maxBatch])");
	if (qbf_transport_periodic(e, dW, ncell, L, h, 0.5 * dt, c->sw)) return cnn_err(-2, "%s", qbf_last_error());
	int r = qbf_surrogate_collision_step(c, e, c->sw, ncell, dt, iscale, oscale, L, mask, c->sw2);
	if (r) return r;
	if (qbf_transport_periodic(e, c->sw2, ncell, L, h, 0.5 * dt, dOut)) return cnn_err(-2, "%s", qbf_last_error());
	return 0;
}

namespace {
struct SurrogateStepCall {
	qbf_cnn* c;
	qbf_engine* e;
	const double* W;
	int ncell;
	double L, h, dt;
	float iscale, oscale;
	int mask;
	double* out;
};
int surrogate_step_call(void* p)
{
	auto* a = (SurrogateStepCall*)p;
	return qbf_surrogate_inhom_step_periodic(a->c, a->e, a->W, a->ncell, a->L, a->h, a->dt, a->iscale, a->oscale,
		a->mask, a->out);
}
} // namespace

int qbf_graph_capture_surrogate_inhom_step_periodic(qbf_cnn* c, qbf_engine* e, const double* dW, int ncell, double L,
	double h, double dt, float iscale, float oscale, int mask, double* dOut, qbf_graph** g)
{
	if (!c || !e || !dW || !dOut || !g) return cnn_err(-1, "qbf_graph_capture_surrogate_inhom_step_periodic: null argument");
	SurrogateStepCall a{c, e, dW, ncell, L, h, dt, iscale, oscale, mask, dOut};
	if (qbf_graph_capture_calls(e, surrogate_step_call, &a, g)) return cnn_err(-2, "%s", qbf_last_error());
	return 0;
}

} // extern "C"
