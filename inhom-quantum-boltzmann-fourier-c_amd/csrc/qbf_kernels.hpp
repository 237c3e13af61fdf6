// qbf_kernels.hpp -- argument blocks shared by the HIP kernels and the host engine.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qbf {

// Pruned 2-D FFT of zero-padded, optionally pre-multiplied N x N Fourier grids:
//   out[o] = FFT2( pad( pre_q .* src ) )        (R_b native order, P = 2N)
// o = (cell * nq + qq) * nslot + slot, src grid = src[cell * nsrc + srcOfSlot[slot]],
// pre = pre[q0 + qq] (none / real N*N / complex N*N).
struct PadFFTArgs {
	void* xs;             // per-workgroup exchange scratch (global), or null for LDS
	const double2* src;
	int nsrc;
	const void* pre;
	int preKind;          // 0 none, 1 real double, 2 complex double2
	int q0, nq;
	int nslot;
	int srcOfSlot[4];
	int ngrid;            // ncell * nq * nslot grids, two per workgroup
	void* out;
	const double2* tw256;
};

// Unpacking of packed real results (paired path)
struct UnpackArgs {
	const void* in;
	void* out;
	int npk, nout, nch;
	int dst[4][2];
};

// I1-type accumulation (I1 for Cc, I1 with FourierI4 tables for I4):
//   out[cell][inst] = FFT2( sum_g psi1_g .* IFFT2( (fa*Fhat + fb) .* FFT2(pad(phi_g .* src)) ) )
struct I1Args {
	void* xs;             // per-workgroup exchange scratch (global), or null for LDS
	const double2* src;   // [ncell][nsrc][N*N]
	int nsrc;
	const void* Fhat;     // [ncell][nFh][P*P] R_b native
	int nFh;
	double fa, fb;
	const double* phi;    // [nterm][N*N]
	const void* psi1n;    // [nterm][P*P] C_i native
	int nterm;
	int ninst;
	int instSrc[8];
	int instFh[8];
	int nchunk;           // terms split into nchunk ranges (partial sums, added by k_combine)
	void* out;            // [ncell][ninst][nchunk][P*P] R_b native
	const double2* tw256;
};

// I2 accumulation: out[cell][inst] = sum_q FFT2(pad(psi1_q .* f)) .* FFT2(pad(psi2_q .* g))
struct I2Args {
	void* xs;             // per-workgroup exchange scratch (global), or null for LDS
	const double2* src;
	int nsrc;
	const double* psi1;   // [nterm][N*N]
	const double* psi2;   // [nterm][N*N]
	const double2* psic;  // [nterm][N*N] psi1 + i psi2 (paired path)
	int nterm;
	int ninst;
	int instF[8];
	int instG[8];
	int nchunk;           // terms split into nchunk ranges (partial sums, added by k_combine)
	void* out;            // [ncell][ninst][nchunk][P*P] R_b native
	const double2* tw256;
};

// I3 accumulation over a range of quadrature terms (the dominant kernel):
//   acc[cell][out][chunk] (+)= sum_{qq in chunk} sum_inst w * F[f] .* K_q[ Shat[g] .* F[h] ],
//   K_q[X] = FFT2( psi2_q .* IFFT2(X) ).
struct I3Args {
	void* xs;             // per-workgroup exchange scratch (global), or null for LDS
	const void* Shat;     // [ncell][nS][P*P] R_b native
	int nS;
	const void* F;        // [ncell][nq][nF][P*P] R_b native
	int nF, nq, q0;
	const void* psi2n;    // [nqTotal][P*P] C_i native
	int nout;
	int ninst[4];
	int instF[4][4], instS[4][4], instH[4][4];
	double instW[4][4];
	int nchunk;
	int accumulate;
	void* acc;            // [ncell][nout][nchunk][P*P] R_b native
	const double2* tw256;
};

// One product term of the physical-domain combination:
//   coef * (X ? ax*X + bx : 1) .* (Y ? sum_{n<nsum} Y[n*ysum] : 1)
struct CTerm {
	const void* X;
	long long xcell;
	const void* Y;
	long long ycell;
	long long ysum;
	int nsum;
	double coef, ax, bx;
};

// Combine terms per output component, IFFT2, extract the N x N block, then
//   out = base + cf * (outScale * value + CB(Wcb, Bext)).
struct CombineArgs {
	void* xs;             // per-workgroup exchange scratch (global), or null for LDS
	const CTerm* terms;   // [nout][maxTerms]
	int maxTerms;
	int nterm[4];
	int nout;
	int outComp[4];       // component index written for each output
	double outScale;
	double2* out;         // [ncell][4][N*N]
	const double2* base;  // [ncell][4][N*N] or null
	const double2* Wcb;   // [ncell][4][N*N] or null (no magnetic rotation)
	const double* Bext;   // [ncell][3] (used when Wcb != null)
	double cf;
	const double2* tw256;
};

} // namespace qbf
