// dgx_mlp.hip -- learned block preconditioner: a 3-layer MLP (BF16 weights, FP32 accumulation on
// MFMA) maps per-element features to an approximate inverse P_e of the element diagonal block
// A_e = dR_e/du_e - M_e/dt_e of the pseudo-time Newton system, in one of two structures:
//   LOWRANK: P_e = sigma_e (diag(d_e) + U_e V_e^T), rank R (= 8);
//   KRON:    P_e = sigma_e sum_{r<R} S_r (x) Q_r, S_r nv x nv (variables), Q_r Np x Np (nodes), R = 4 or 8,
//            which follows the (variables) x (nodes) Kronecker structure of the DG block;
// sigma_e = 1 / (mean(w detJ)/dt_e + 0.25 h_e lambda_e) is analytic, so the network outputs stay O(1)
// along the CFL ramp.
// Training data are the solver's own blocks captured along its Newton iterations; the loss is
//   L = (1/B) sum_e ||P_e A_e - I||_F^2 / nb,
// minimised with Adam.  Forward, backward and the loss gradient are hand-written kernels; the GEMMs use
// v_mfma_f32_16x16x16_bf16 (one 16x16 output tile per wavefront, K in steps of 16).
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "dgx_device.h"
#include "dgx_internal.h"

namespace dgx {

using bf16x4 = __attribute__((__vector_size__(4 * sizeof(short)))) short;
using f32x4 = __attribute__((__vector_size__(4 * sizeof(float)))) float;

__device__ inline short f2bf(float x)
{
	// round-to-nearest-even float -> bf16 bits
	unsigned u = __float_as_uint(x);
	if ((u & 0x7f800000u) == 0x7f800000u) return (short)(u >> 16);  // inf/nan
	u += 0x7fffu + ((u >> 16) & 1u);
	return (short)(u >> 16);
}
__device__ inline float bf2f(short b) { return __uint_as_float(((unsigned)(unsigned short)b) << 16); }

// C[M x N] (+)= op(A)[M x K] * op(B)[K x N], A/B given as FP32 and rounded to bf16 on load.
// A element (i,k) at A[i*sa0 + k*sa1]; B element (k,j) at B[k*sb0 + j*sb1]; C row-major ldc.
// One wavefront per 16x16 output tile; out-of-range rows/cols/ks read as zero.
__global__ void __launch_bounds__(64) k_gemm_bf16(const float* __restrict__ A, long sa0, long sa1,
                                                  const float* __restrict__ B, long sb0, long sb1, float* C, long ldc,
                                                  int M, int N, int K, int accumulate, const float* __restrict__ bias,
                                                  int act)
{
	const int l = threadIdx.x;
	const int i0 = blockIdx.y * 16, j0 = blockIdx.x * 16;
	const int row = i0 + (l & 15), col = j0 + (l & 15), kq = (l >> 4) * 4;
	f32x4 acc = { 0.f, 0.f, 0.f, 0.f };
	int k0 = 0;
	// four k-steps of operands in flight before their MFMAs (the long-K products are load-latency bound;
	// eight measured slower); the accumulation order is the one of the single-step loop below
	for (; k0 + 64 <= K; k0 += 64) {
		bf16x4 a[4], b[4];
#pragma unroll
		for (int u = 0; u < 4; ++u) {
#pragma unroll
			for (int t = 0; t < 4; ++t) {
				const int k = k0 + 16 * u + kq + t;
				a[u][t] = f2bf(row < M ? A[(long)row * sa0 + (long)k * sa1] : 0.f);
				b[u][t] = f2bf(col < N ? B[(long)k * sb0 + (long)col * sb1] : 0.f);
			}
		}
#pragma unroll
		for (int u = 0; u < 4; ++u) acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[u], b[u], acc, 0, 0, 0);
	}
	for (; k0 < K; k0 += 16) {
		bf16x4 a, b;
#pragma unroll
		for (int t = 0; t < 4; ++t) {
			const int k = k0 + kq + t;
			const float av = (row < M && k < K) ? A[(long)row * sa0 + (long)k * sa1] : 0.f;
			const float bv = (col < N && k < K) ? B[(long)k * sb0 + (long)col * sb1] : 0.f;
			a[t] = f2bf(av);
			b[t] = f2bf(bv);
		}
		acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
	}
#pragma unroll
	for (int t = 0; t < 4; ++t) {
		const int r = i0 + (l >> 4) * 4 + t, c = j0 + (l & 15);
		if (r < M && c < N) {
			float v = acc[t];
			if (accumulate) v += C[(long)r * ldc + c];
			if (bias) v += bias[c];
			if (act == 1) v = tanhf(v);
			C[(long)r * ldc + c] = v;
		}
	}
}

static hipError_t gemm(hipStream_t s, const float* A, long sa0, long sa1, const float* B, long sb0, long sb1, float* C,
                       long ldc, int M, int N, int K, int accumulate, const float* bias, int act)
{
	dim3 grid((N + 15) / 16, (M + 15) / 16);
	hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(64), 0, s, A, sa0, sa1, B, sb0, sb1, C, ldc, M, N, K, accumulate,
	                   bias, act);
	return hipGetLastError();
}

// ------------------------------------------------------------------------------- loss and its gradient
// Per sample e: out = [d (nb) | U (nb x R, row-major) | V (nb x R)], A_e (nb x nb row-major), sigma_e.
// P = sigma (diag(d) + U V^T), Rm = P A - I, L_e = ||Rm||_F^2 / nb.
// dL/dP = 2 Rm A^T / nb; with G = sigma dL/dP: dd = diag(G), dU = G V, dV = G^T U  (times gscale).
template <int NB, int RK>
__global__ void __launch_bounds__(256) k_block_loss(const float* __restrict__ out, const float* __restrict__ sigma,
                                                    const double* __restrict__ A, float* grad, double* loss_partial,
                                                    float gscale)
{
	constexpr int NO = NB + 2 * NB * RK;
	__shared__ float sA[NB * NB], sP[NB * NB], sR[NB * NB];
	__shared__ float sD[NB], sU[NB * RK], sV[NB * RK];
	__shared__ double red[256];
	const int e = blockIdx.x, t = threadIdx.x;
	const float sg = sigma ? sigma[e] : 1.f;
	const float* o = out + (size_t)e * NO;
	for (int i = t; i < NB * NB; i += blockDim.x) sA[i] = (float)A[(size_t)e * NB * NB + i];
	for (int i = t; i < NB; i += blockDim.x) sD[i] = o[i];
	for (int i = t; i < NB * RK; i += blockDim.x) { sU[i] = o[NB + i]; sV[i] = o[NB + NB * RK + i]; }
	__syncthreads();
	for (int ij = t; ij < NB * NB; ij += blockDim.x) {
		const int i = ij / NB, j = ij % NB;
		float p = (i == j) ? sD[i] : 0.f;
		for (int r = 0; r < RK; ++r) p += sU[i * RK + r] * sV[j * RK + r];
		sP[ij] = sg * p;
	}
	__syncthreads();
	double lsum = 0.0;
	for (int ij = t; ij < NB * NB; ij += blockDim.x) {
		const int i = ij / NB, j = ij % NB;
		float acc = (i == j) ? -1.f : 0.f;
		for (int k = 0; k < NB; ++k) acc += sP[i * NB + k] * sA[k * NB + j];
		sR[ij] = acc;
		lsum += (double)acc * acc;
	}
	red[t] = lsum;
	__syncthreads();
	for (int off = 128; off > 0; off >>= 1) {
		if (t < off) red[t] += red[t + off];
		__syncthreads();
	}
	if (t == 0) loss_partial[e] = red[0] / NB;
	if (!grad) return;
	// G = sigma * 2 Rm A^T / nb * gscale  (stored in sP, which is no longer needed)
	for (int ij = t; ij < NB * NB; ij += blockDim.x) {
		const int i = ij / NB, j = ij % NB;
		float acc = 0.f;
		for (int k = 0; k < NB; ++k) acc += sR[i * NB + k] * sA[j * NB + k];
		sP[ij] = acc * (2.0f / NB) * gscale * sg;
	}
	__syncthreads();
	float* g = grad + (size_t)e * NO;
	for (int i = t; i < NB; i += blockDim.x) g[i] = sP[i * NB + i];
	for (int ir = t; ir < NB * RK; ir += blockDim.x) {
		const int i = ir / RK, r = ir % RK;
		float du = 0.f, dv = 0.f;
		for (int j = 0; j < NB; ++j) {
			du += sP[i * NB + j] * sV[j * RK + r];
			dv += sP[j * NB + i] * sU[j * RK + r];
		}
		g[NB + ir] = du;
		g[NB + NB * RK + ir] = dv;
	}
}

// z_e = sigma_e (d_e * v_e + U_e (V_e^T v_e))   (FP64 vectors, FP32 factors)
template <int NB, int RK>
__global__ void __launch_bounds__(64) k_lowrank_apply(const float* __restrict__ out, const float* __restrict__ sigma,
                                                      const double* __restrict__ v, double* z, int E)
{
	constexpr int NO = NB + 2 * NB * RK;
	__shared__ double sv[NB], sw[RK];
	const int e = blockIdx.x, t = threadIdx.x;
	const float* o = out + (size_t)e * NO;
	const double sg = sigma[e];
	for (int i = t; i < NB; i += blockDim.x) sv[i] = v[(size_t)e * NB + i];
	__syncthreads();
	for (int r = t; r < RK; r += blockDim.x) {
		double s = 0.0;
		for (int j = 0; j < NB; ++j) s += (double)o[NB + NB * RK + j * RK + r] * sv[j];
		sw[r] = s;
	}
	__syncthreads();
	for (int i = t; i < NB; i += blockDim.x) {
		double s = (double)o[i] * sv[i];
		for (int r = 0; r < RK; ++r) s += (double)o[NB + i * RK + r] * sw[r];
		z[(size_t)e * NB + i] = sg * s;
	}
}

// KRON structure.  out = [S_0 .. S_{RK-1} (NV x NV each) | Q_0 .. Q_{RK-1} (NPN x NPN each)], row-major;
// P[(a,i),(b,j)] = sigma sum_r S_r[a,b] Q_r[i,j] (rows var-major like the state).
// With G = sigma dL/dP: dS_r[a,b] = sum_ij G[(a,i),(b,j)] Q_r[i,j], dQ_r[i,j] = sum_ab G[(a,i),(b,j)] S_r[a,b].
template <int NV, int NPN, int RK>
__global__ void __launch_bounds__(256) k_block_loss_kron(const float* __restrict__ out,
                                                         const float* __restrict__ sigma,
                                                         const double* __restrict__ A, float* grad,
                                                         double* loss_partial, float gscale)
{
	constexpr int NB = NV * NPN, NS = NV * NV, NQ = NPN * NPN, NO = RK * (NS + NQ);
	// 64 x 64 blocks: the products run as 4 x 4 output tiles per thread (256 threads); other sizes one output
	// per thread-iteration; rows padded to NB + 1 floats against LDS bank conflicts; every output is summed
	// over k in ascending order
	constexpr bool TILE = (NB == 64);
	constexpr int LD = NB + 1;
	__shared__ float sA[NB * LD], sP[NB * LD], sR[NB * LD];
	__shared__ float sO[NO];
	__shared__ double red[256];
	const int e = blockIdx.x, t = threadIdx.x;
	const int ti = t / 16, tj = t % 16;   // output rows 4 ti .. 4 ti + 3, columns 4 tj .. 4 tj + 3
	const float sg = sigma ? sigma[e] : 1.f;
	const float* o = out + (size_t)e * NO;
	for (int i = t; i < NB * NB; i += blockDim.x) sA[(i / NB) * LD + i % NB] = (float)A[(size_t)e * NB * NB + i];
	for (int i = t; i < NO; i += blockDim.x) sO[i] = o[i];
	__syncthreads();
	const float* S = sO;
	const float* Q = sO + RK * NS;
	for (int ij = t; ij < NB * NB; ij += blockDim.x) {
		const int row = ij / NB, col = ij % NB;
		const int a = row / NPN, i = row % NPN, b = col / NPN, j = col % NPN;
		float p = 0.f;
		for (int r = 0; r < RK; ++r) p += S[r * NS + a * NV + b] * Q[r * NQ + i * NPN + j];
		sP[row * LD + col] = sg * p;
	}
	__syncthreads();
	// R = P A - I
	double lsum = 0.0;
	if constexpr (!TILE) {
		for (int ij = t; ij < NB * NB; ij += blockDim.x) {
			const int i = ij / NB, j = ij % NB;
			float acc = (i == j) ? -1.f : 0.f;
			for (int k = 0; k < NB; ++k) acc += sP[i * LD + k] * sA[k * LD + j];
			sR[i * LD + j] = acc;
			lsum += (double)acc * acc;
		}
	} else {
		float acc[4][4];
#pragma unroll
		for (int r = 0; r < 4; ++r)
#pragma unroll
			for (int c = 0; c < 4; ++c) acc[r][c] = (4 * ti + r == 4 * tj + c) ? -1.f : 0.f;
		for (int k = 0; k < NB; ++k) {
			float pa[4], ab[4];
#pragma unroll
			for (int r = 0; r < 4; ++r) pa[r] = sP[(4 * ti + r) * LD + k];
#pragma unroll
			for (int c = 0; c < 4; ++c) ab[c] = sA[k * LD + 4 * tj + c];
#pragma unroll
			for (int r = 0; r < 4; ++r)
#pragma unroll
				for (int c = 0; c < 4; ++c) acc[r][c] += pa[r] * ab[c];
		}
#pragma unroll
		for (int r = 0; r < 4; ++r)
#pragma unroll
			for (int c = 0; c < 4; ++c) {
				sR[(4 * ti + r) * LD + 4 * tj + c] = acc[r][c];
				lsum += (double)acc[r][c] * acc[r][c];
			}
	}
	red[t] = lsum;
	__syncthreads();
	for (int off = 128; off > 0; off >>= 1) {
		if (t < off) red[t] += red[t + off];
		__syncthreads();
	}
	if (t == 0) loss_partial[e] = red[0] / NB;
	if (!grad) return;
	// G = sigma * 2 R A^T / nb * gscale  (into sP, no longer needed)
	if constexpr (!TILE) {
		for (int ij = t; ij < NB * NB; ij += blockDim.x) {
			const int i = ij / NB, j = ij % NB;
			float acc = 0.f;
			for (int k = 0; k < NB; ++k) acc += sR[i * LD + k] * sA[j * LD + k];
			sP[i * LD + j] = acc * (2.0f / NB) * gscale * sg;
		}
	} else {
		float acc[4][4];
#pragma unroll
		for (int r = 0; r < 4; ++r)
#pragma unroll
			for (int c = 0; c < 4; ++c) acc[r][c] = 0.f;
		for (int k = 0; k < NB; ++k) {
			float ra[4], aa[4];
#pragma unroll
			for (int r = 0; r < 4; ++r) ra[r] = sR[(4 * ti + r) * LD + k];
#pragma unroll
			for (int c = 0; c < 4; ++c) aa[c] = sA[(4 * tj + c) * LD + k];
#pragma unroll
			for (int r = 0; r < 4; ++r)
#pragma unroll
				for (int c = 0; c < 4; ++c) acc[r][c] += ra[r] * aa[c];
		}
#pragma unroll
		for (int r = 0; r < 4; ++r)
#pragma unroll
			for (int c = 0; c < 4; ++c) sP[(4 * ti + r) * LD + 4 * tj + c] = acc[r][c] * (2.0f / NB) * gscale * sg;
	}
	__syncthreads();
	float* g = grad + (size_t)e * NO;
	for (int x = t; x < RK * NS; x += blockDim.x) {
		const int r = x / NS, a = (x % NS) / NV, b = x % NV;
		float acc = 0.f;
		for (int i = 0; i < NPN; ++i)
			for (int j = 0; j < NPN; ++j) acc += sP[(a * NPN + i) * LD + b * NPN + j] * Q[r * NQ + i * NPN + j];
		g[x] = acc;
	}
	for (int x = t; x < RK * NQ; x += blockDim.x) {
		const int r = x / NQ, i = (x % NQ) / NPN, j = x % NPN;
		float acc = 0.f;
		for (int a = 0; a < NV; ++a)
			for (int b = 0; b < NV; ++b) acc += sP[(a * NPN + i) * LD + b * NPN + j] * S[r * NS + a * NV + b];
		g[RK * NS + x] = acc;
	}
}

// z_e[a,i] = sigma_e sum_r sum_b S_r[a,b] sum_j Q_r[i,j] v_e[b,j]   (FP64 vectors, FP32 factors)
template <int NV, int NPN, int RK>
__global__ void __launch_bounds__(64) k_kron_apply(const float* __restrict__ out, const float* __restrict__ sigma,
                                                   const double* __restrict__ v, double* z, int E)
{
	constexpr int NB = NV * NPN, NS = NV * NV, NQ = NPN * NPN, NO = RK * (NS + NQ);
	__shared__ float sO[NO];
	__shared__ double sv[NB], sW[RK * NB];
	const int e = blockIdx.x, t = threadIdx.x;
	const float* o = out + (size_t)e * NO;
	for (int i = t; i < NO; i += blockDim.x) sO[i] = o[i];
	for (int i = t; i < NB; i += blockDim.x) sv[i] = v[(size_t)e * NB + i];
	__syncthreads();
	const float* S = sO;
	const float* Q = sO + RK * NS;
	for (int x = t; x < RK * NB; x += blockDim.x) {   // W_r[b][i] = sum_j Q_r[i,j] v[b,j]
		const int r = x / NB, b = (x % NB) / NPN, i = x % NPN;
		double acc = 0.0;
		for (int j = 0; j < NPN; ++j) acc += (double)Q[r * NQ + i * NPN + j] * sv[b * NPN + j];
		sW[x] = acc;
	}
	__syncthreads();
	const double sg = sigma[e];
	for (int x = t; x < NB; x += blockDim.x) {
		const int a = x / NPN, i = x % NPN;
		double acc = 0.0;
		for (int r = 0; r < RK; ++r)
			for (int b = 0; b < NV; ++b) acc += (double)S[r * NS + a * NV + b] * sW[r * NB + b * NPN + i];
		z[(size_t)e * NB + x] = sg * acc;
	}
}

// features [u / u_ref (nb) | detJ / mean (Np) | log10(h inv_dt / lambda) | 0 pad] and sigma_e
template <int D, int N>
__global__ void __launch_bounds__(64) k_features(const double* __restrict__ u, const double* __restrict__ detJ,
                                                 const Ops1D<double>* ops, const double* __restrict__ hmin,
                                                 const double* __restrict__ inv_dt, int E, float* feat, int nf,
                                                 double uref0, double uref1, double uref2, float* sigma)
{
	constexpr int Np = ipow(N, D), NV = D + 2, nb = NV * Np;
	__shared__ double sWJ[Np], sWs[Np];
	const int e = blockIdx.x, t = threadIdx.x;
	for (int q = t; q < Np; q += blockDim.x) {
		const int c0 = q % N, c1 = (q / N) % N, c2 = (D == 3) ? q / (N * N) : 0;
		double w = ops->w[c0] * ops->w[c1];
		if (D == 3) w *= ops->w[c2];
		sWJ[q] = w * detJ[(size_t)e * Np + q];
		const double* ue = u + (size_t)e * nb;
		const double rho = ue[q], ri = 1.0 / rho;
		double V2 = 0.0;
		for (int k = 0; k < D; ++k) { const double w2 = ue[(1 + k) * Np + q] * ri; V2 += w2 * w2; }
		const double p = kGM1 * (ue[(D + 1) * Np + q] - 0.5 * rho * V2);
		sWs[q] = sqrt(V2) + sqrt(fmax(kGamma * p * ri, 0.0));
	}
	__syncthreads();
	__shared__ double mj, mdj, ws;
	if (t == 0) {
		double a = 0.0, b = 0.0, m = 0.0;
		for (int q = 0; q < Np; ++q) { a += sWJ[q]; b += detJ[(size_t)e * Np + q]; m = fmax(m, sWs[q]); }
		mj = a / Np;
		mdj = b / Np;
		ws = m;
		sigma[e] = (float)(1.0 / (mj * inv_dt[e] + 0.25 * hmin[e] * m));
	}
	__syncthreads();
	float* f = feat + (size_t)e * nf;
	for (int i = t; i < nb; i += blockDim.x) {
		const int v = i / Np;
		const double ref = (v == 0) ? uref0 : (v == NV - 1 ? uref2 : uref1);
		f[i] = (float)(u[(size_t)e * nb + i] / ref);
	}
	for (int i = t; i < Np; i += blockDim.x) f[nb + i] = (float)(detJ[(size_t)e * Np + i] / mdj);
	if (t == 0) f[nb + Np] = (float)log10(hmin[e] * inv_dt[e] / ws + 1e-30);
	for (int i = nb + Np + 1 + t; i < nf; i += blockDim.x) f[i] = 0.f;
}

__global__ void k_tanh_backward(const float* __restrict__ h, float* dh, size_t n)
{
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		dh[i] *= (1.f - h[i] * h[i]);
}
// out[j] = sum_b g[b][j]: 64 columns per workgroup, the rows split over 16 slices whose partial sums are added
// in a fixed order (deterministic); a row of 64 columns is one coalesced load per slice
__global__ void __launch_bounds__(1024) k_colsum(const float* __restrict__ g, int B, int N, float* out)
{
	__shared__ float part[16][64];
	const int c = threadIdx.x & 63, sl = threadIdx.x >> 6;
	const int j = blockIdx.x * 64 + c;
	float s = 0.f;
	if (j < N)
		for (int b = sl; b < B; b += 16) s += g[(size_t)b * N + j];
	part[sl][c] = s;
	__syncthreads();
	if (sl == 0 && j < N) {
		float t = part[0][c];
#pragma unroll
		for (int k = 1; k < 16; ++k) t += part[k][c];
		out[j] = t;
	}
}
__global__ void k_adam(float* w, const float* __restrict__ g, float* m, float* v, size_t n, float lr, float b1,
                       float b2, float c1, float c2, float eps)
{
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
		const float gi = g[i];
		m[i] = b1 * m[i] + (1.f - b1) * gi;
		v[i] = b2 * v[i] + (1.f - b2) * gi * gi;
		w[i] -= lr * (m[i] / c1) / (sqrtf(v[i] / c2) + eps);
	}
}
// the forward pass always sees the BF16 values of the FP32 master weights
__global__ void k_round_bf16(const float* __restrict__ w, float* wb, size_t n)
{
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		wb[i] = bf2f(f2bf(w[i]));
}

static int gridn(size_t n) { return (int)((n + 255) / 256 < 8192 ? (n + 255) / 256 : 8192); }

}  // namespace dgx

// ------------------------------------------------------------------------------- the model
struct dgx_mlp {
	int kind = 0;
	int nf = 0, nhid = 0, nout = 0, nb = 0, rank = 0, Np = 0;
	float* w = nullptr;   // FP32 master parameters W1,b1,W2,b2,W3,b3
	float* wb = nullptr;  // BF16-rounded copy used by the forward pass
	float* gw = nullptr;
	float* am = nullptr;
	float* av = nullptr;
	size_t nparam = 0;
	size_t off[6] = { 0, 0, 0, 0, 0, 0 };
	int step = 0;
	int cap = 0;
	float *h1 = nullptr, *h2 = nullptr, *o = nullptr, *go = nullptr, *gh2 = nullptr, *gh1 = nullptr;
	double* lpart = nullptr;
	hipStream_t s = nullptr;
};

namespace dgx {

static void mlp_free_work(dgx_mlp* m)
{
	for (float* p : { m->h1, m->h2, m->o, m->go, m->gh2, m->gh1 })
		if (p) (void)hipFree(p);
	if (m->lpart) (void)hipFree(m->lpart);
	m->h1 = m->h2 = m->o = m->go = m->gh2 = m->gh1 = nullptr;
	m->lpart = nullptr;
	m->cap = 0;
}

static hipError_t mlp_reserve(dgx_mlp* m, int B)
{
	if (B <= m->cap) return hipSuccess;
	mlp_free_work(m);
	DGX_HIP_CHECK(hipMalloc(&m->h1, sizeof(float) * (size_t)B * m->nhid));
	DGX_HIP_CHECK(hipMalloc(&m->h2, sizeof(float) * (size_t)B * m->nhid));
	DGX_HIP_CHECK(hipMalloc(&m->o, sizeof(float) * (size_t)B * m->nout));
	DGX_HIP_CHECK(hipMalloc(&m->go, sizeof(float) * (size_t)B * m->nout));
	DGX_HIP_CHECK(hipMalloc(&m->gh2, sizeof(float) * (size_t)B * m->nhid));
	DGX_HIP_CHECK(hipMalloc(&m->gh1, sizeof(float) * (size_t)B * m->nhid));
	DGX_HIP_CHECK(hipMalloc(&m->lpart, sizeof(double) * (size_t)B));
	m->cap = B;
	return hipSuccess;
}

static hipError_t mlp_forward(dgx_mlp* m, const float* X, int B)
{
	DGX_HIP_CHECK(mlp_reserve(m, B));
	const float *W1 = m->wb + m->off[0], *b1 = m->wb + m->off[1], *W2 = m->wb + m->off[2], *b2 = m->wb + m->off[3],
	            *W3 = m->wb + m->off[4], *b3 = m->wb + m->off[5];
	DGX_HIP_CHECK(gemm(m->s, X, m->nf, 1, W1, 1, m->nf, m->h1, m->nhid, B, m->nhid, m->nf, 0, b1, 1));
	DGX_HIP_CHECK(gemm(m->s, m->h1, m->nhid, 1, W2, 1, m->nhid, m->h2, m->nhid, B, m->nhid, m->nhid, 0, b2, 1));
	return gemm(m->s, m->h2, m->nhid, 1, W3, 1, m->nhid, m->o, m->nout, B, m->nout, m->nhid, 0, b3, 0);
}

static hipError_t mlp_backward(dgx_mlp* m, const float* X, int B)
{
	const float *W2 = m->wb + m->off[2], *W3 = m->wb + m->off[4];
	float *gW1 = m->gw + m->off[0], *gb1 = m->gw + m->off[1], *gW2 = m->gw + m->off[2], *gb2 = m->gw + m->off[3],
	      *gW3 = m->gw + m->off[4], *gb3 = m->gw + m->off[5];
	DGX_HIP_CHECK(gemm(m->s, m->go, 1, m->nout, m->h2, m->nhid, 1, gW3, m->nhid, m->nout, m->nhid, B, 0, nullptr, 0));
	hipLaunchKernelGGL(k_colsum, dim3((m->nout + 63) / 64), dim3(1024), 0, m->s, m->go, B, m->nout, gb3);
	DGX_HIP_CHECK(gemm(m->s, m->go, m->nout, 1, W3, m->nhid, 1, m->gh2, m->nhid, B, m->nhid, m->nout, 0, nullptr, 0));
	hipLaunchKernelGGL(k_tanh_backward, dim3(gridn((size_t)B * m->nhid)), dim3(256), 0, m->s, m->h2, m->gh2,
	                   (size_t)B * m->nhid);
	DGX_HIP_CHECK(gemm(m->s, m->gh2, 1, m->nhid, m->h1, m->nhid, 1, gW2, m->nhid, m->nhid, m->nhid, B, 0, nullptr, 0));
	hipLaunchKernelGGL(k_colsum, dim3((m->nhid + 63) / 64), dim3(1024), 0, m->s, m->gh2, B, m->nhid, gb2);
	DGX_HIP_CHECK(gemm(m->s, m->gh2, m->nhid, 1, W2, m->nhid, 1, m->gh1, m->nhid, B, m->nhid, m->nhid, 0, nullptr, 0));
	hipLaunchKernelGGL(k_tanh_backward, dim3(gridn((size_t)B * m->nhid)), dim3(256), 0, m->s, m->h1, m->gh1,
	                   (size_t)B * m->nhid);
	DGX_HIP_CHECK(gemm(m->s, m->gh1, 1, m->nhid, X, m->nf, 1, gW1, m->nf, m->nhid, m->nf, B, 0, nullptr, 0));
	hipLaunchKernelGGL(k_colsum, dim3((m->nhid + 63) / 64), dim3(1024), 0, m->s, m->gh1, B, m->nhid, gb1);
	return hipGetLastError();
}

static hipError_t mlp_loss(dgx_mlp* m, const float* sigma, const double* A, int B, double* loss, float gscale,
                           bool want_grad)
{
	float* g = want_grad ? m->go : nullptr;
	if (m->kind == DGX_MLP_LOWRANK && m->nb == 64 && m->rank == 8)
		hipLaunchKernelGGL((k_block_loss<64, 8>), dim3(B), dim3(256), 0, m->s, m->o, sigma, A, g, m->lpart, gscale);
	else if (m->kind == DGX_MLP_LOWRANK && m->nb == 36 && m->rank == 8)
		hipLaunchKernelGGL((k_block_loss<36, 8>), dim3(B), dim3(256), 0, m->s, m->o, sigma, A, g, m->lpart, gscale);
#define DGX_KRON_LOSS(NPN, RK)                                                                                  \
	else if (m->kind == DGX_MLP_KRON && m->Np == NPN && m->rank == RK)                                          \
		hipLaunchKernelGGL((k_block_loss_kron<4, NPN, RK>), dim3(B), dim3(256), 0, m->s, m->o, sigma, A, g, m->lpart, \
		                   gscale);
	DGX_KRON_LOSS(16, 4) DGX_KRON_LOSS(9, 4) DGX_KRON_LOSS(16, 8) DGX_KRON_LOSS(9, 8)
#undef DGX_KRON_LOSS
	else
		return hipErrorInvalidValue;
	DGX_HIP_CHECK(hipGetLastError());
	if (loss) {
		std::vector<double> lp(B);
		DGX_HIP_CHECK(hipMemcpyAsync(lp.data(), m->lpart, sizeof(double) * B, hipMemcpyDeviceToHost, m->s));
		DGX_HIP_CHECK(hipStreamSynchronize(m->s));
		double sum = 0.0;
		for (double x : lp) sum += x;
		*loss = sum / B;
	}
	return hipSuccess;
}

static hipError_t mlp_sync_bf16(dgx_mlp* m)
{
	hipLaunchKernelGGL(k_round_bf16, dim3(gridn(m->nparam)), dim3(256), 0, m->s, m->w, m->wb, m->nparam);
	return hipGetLastError();
}

struct PrecondLearned : Precond {
	dgx_mlp* mlp = nullptr;
	float* feat = nullptr;
	float* sigma = nullptr;
	~PrecondLearned() override
	{
		if (feat) (void)hipFree(feat);
		if (sigma) (void)hipFree(sigma);
	}
	hipError_t setup(Ctx& c, const double* u, const double* inv_dt) override;
	hipError_t apply(Ctx& c, const double* v, double* z) override
	{
		if (mlp->kind == DGX_MLP_KRON) {
#define DGX_KRON_APPLY(NPN, RK)                                                                                     \
			if (mlp->Np == NPN && mlp->rank == RK) {                                                                \
				hipLaunchKernelGGL((k_kron_apply<4, NPN, RK>), dim3(c.E), dim3(64), 0, c.stream, mlp->o, sigma, v, z, \
				                   c.E);                                                                            \
				return hipGetLastError();                                                                          \
			}
			DGX_KRON_APPLY(16, 4) DGX_KRON_APPLY(9, 4) DGX_KRON_APPLY(16, 8) DGX_KRON_APPLY(9, 8)
#undef DGX_KRON_APPLY
			return hipErrorInvalidValue;
		}
		if (mlp->nb == 64 && mlp->rank == 8)
			hipLaunchKernelGGL((k_lowrank_apply<64, 8>), dim3(c.E), dim3(64), 0, c.stream, mlp->o, sigma, v, z, c.E);
		else if (mlp->nb == 36 && mlp->rank == 8)
			hipLaunchKernelGGL((k_lowrank_apply<36, 8>), dim3(c.E), dim3(64), 0, c.stream, mlp->o, sigma, v, z, c.E);
		else
			return hipErrorInvalidValue;
		return hipGetLastError();
	}
};

hipError_t launch_features(Ctx& c, const double* u, const double* inv_dt, float* feat, int nf, float* sigma)
{
	if (c.d != 2 || c.nv != 4) return hipErrorInvalidValue;
	const double* bp = c.cfg.bc_prim;
	const double rho = bp[0] > 0 ? bp[0] : 1.0;
	const double vel = std::sqrt(bp[1] * bp[1] + bp[2] * bp[2]);
	const double uref1 = rho * (vel > 0 ? vel : 1.0);
	const double pres = bp[3] > 0 ? bp[3] : 1.0;
	const double uref2 = pres / kGM1 + 0.5 * rho * vel * vel;
#define DGX_FEAT(NN)                                                                                              \
	if (c.N == NN) {                                                                                              \
		hipLaunchKernelGGL((k_features<2, NN>), dim3(c.E), dim3(64), 0, c.stream, u, c.detJ_vc, c.ops64, c.hmin,   \
		                   inv_dt, c.E, feat, nf, rho, uref1, uref2, sigma);                                       \
		return hipGetLastError();                                                                                 \
	}
	DGX_FEAT(2) DGX_FEAT(3) DGX_FEAT(4) DGX_FEAT(5) DGX_FEAT(6)
#undef DGX_FEAT
	return hipErrorInvalidValue;
}

hipError_t PrecondLearned::setup(Ctx& c, const double* u, const double* inv_dt)
{
	if (!feat) {
		DGX_HIP_CHECK(hipMalloc(&feat, sizeof(float) * (size_t)c.E * mlp->nf));
		DGX_HIP_CHECK(hipMalloc(&sigma, sizeof(float) * (size_t)c.E));
	}
	// inference runs on the solver's stream; the model keeps its own stream for training calls
	const hipStream_t saved = mlp->s;
	mlp->s = c.stream;
	hipError_t err = launch_features(c, u, inv_dt, feat, mlp->nf, sigma);
	if (err == hipSuccess) err = mlp_forward(mlp, feat, c.E);
	mlp->s = saved;
	return err;
}

Precond* make_learned_precond(Ctx& c)
{
	if (!c.mlp || c.mlp->nb != c.Np * c.nv || c.mlp->Np != c.Np) return nullptr;
	PrecondLearned* p = new PrecondLearned();
	p->mlp = c.mlp;
	return p;
}

static uint64_t splitmix64(uint64_t& s)
{
	uint64_t z = (s += 0x9E3779B97F4A7C15ull);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

}  // namespace dgx

extern "C" {

dgx_status dgx_mlp_create(dgx_mlp** out, int kind, int nb, int Np, int hidden, int rank, unsigned long long seed)
{
	if (!out || hidden < 16 || hidden > 1024) return DGX_ERR_ARG;
	// 2-D Euler blocks: nv = 4 variables, p = 2 (Np = 9) or p = 3 (Np = 16)
	if (!((nb == 64 && Np == 16) || (nb == 36 && Np == 9))) return DGX_ERR_ARG;
	if (!((kind == DGX_MLP_LOWRANK && rank == 8) || (kind == DGX_MLP_KRON && (rank == 4 || rank == 8))))
		return DGX_ERR_ARG;
	const int nv = nb / Np;
	dgx_mlp* m = new dgx_mlp();
	m->kind = kind;
	m->nb = nb;
	m->Np = Np;
	m->rank = rank;
	m->nhid = hidden;
	m->nf = ((nb + Np + 1) + 15) / 16 * 16;
	m->nout = (kind == DGX_MLP_LOWRANK) ? nb + 2 * nb * rank : rank * (nv * nv + Np * Np);
	const size_t sizes[6] = { (size_t)m->nhid * m->nf, (size_t)m->nhid, (size_t)m->nhid * m->nhid, (size_t)m->nhid,
	                          (size_t)m->nout * m->nhid, (size_t)m->nout };
	size_t acc = 0;
	for (int i = 0; i < 6; ++i) { m->off[i] = acc; acc += sizes[i]; }
	m->nparam = acc;
	std::vector<float> w(acc, 0.f);
	uint64_t st = seed ? seed : 20261202ull;
	auto uni = [&](float a) { return (float)(((dgx::splitmix64(st) >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0) * a; };
	const int fan[3][2] = { { m->nf, m->nhid }, { m->nhid, m->nhid }, { m->nhid, m->nout } };
	for (int l = 0; l < 3; ++l) {
		const float a = std::sqrt(6.0f / (fan[l][0] + fan[l][1])) * (l == 2 ? 0.1f : 1.0f);
		for (size_t i = 0; i < sizes[2 * l]; ++i) w[m->off[2 * l] + i] = uni(a);
	}
	// start from the (negative) scaled identity P = -sigma I
	if (kind == DGX_MLP_LOWRANK) {
		for (int i = 0; i < nb; ++i) w[m->off[5] + i] = -1.0f;
	} else {
		for (int a = 0; a < nv; ++a) w[m->off[5] + a * nv + a] = -1.0f;                         // S_0 = -I
		for (int i = 0; i < Np; ++i) w[m->off[5] + rank * nv * nv + i * Np + i] = 1.0f;          // Q_0 = I
	}
	if (hipMalloc(&m->w, sizeof(float) * acc) != hipSuccess || hipMalloc(&m->wb, sizeof(float) * acc) != hipSuccess ||
	    hipMalloc(&m->gw, sizeof(float) * acc) != hipSuccess || hipMalloc(&m->am, sizeof(float) * acc) != hipSuccess ||
	    hipMalloc(&m->av, sizeof(float) * acc) != hipSuccess) {
		dgx_mlp_destroy(m);
		return DGX_ERR_HIP;
	}
	std::vector<float> z(acc, 0.f);
	if (hipMemcpy(m->w, w.data(), sizeof(float) * acc, hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(m->am, z.data(), sizeof(float) * acc, hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(m->av, z.data(), sizeof(float) * acc, hipMemcpyHostToDevice) != hipSuccess ||
	    dgx::mlp_sync_bf16(m) != hipSuccess || hipStreamSynchronize(m->s) != hipSuccess) {
		dgx_mlp_destroy(m);
		return DGX_ERR_HIP;
	}
	*out = m;
	return DGX_OK;
}

dgx_status dgx_mlp_destroy(dgx_mlp* m)
{
	if (!m) return DGX_OK;
	dgx::mlp_free_work(m);
	for (float* p : { m->w, m->wb, m->gw, m->am, m->av })
		if (p) (void)hipFree(p);
	delete m;
	return DGX_OK;
}

dgx_status dgx_mlp_sizes(const dgx_mlp* m, int* nf, int* nout, long long* nparam)
{
	if (!m) return DGX_ERR_ARG;
	if (nf) *nf = m->nf;
	if (nout) *nout = m->nout;
	if (nparam) *nparam = (long long)m->nparam;
	return DGX_OK;
}

dgx_status dgx_mlp_get_params(dgx_mlp* m, float* host)
{
	if (!m || !host) return DGX_ERR_ARG;
	// on the model's stream: the training kernels that update m->w run there
	if (hipMemcpyAsync(host, m->w, sizeof(float) * m->nparam, hipMemcpyDeviceToHost, m->s) != hipSuccess ||
	    hipStreamSynchronize(m->s) != hipSuccess)
		return DGX_ERR_HIP;
	return DGX_OK;
}

dgx_status dgx_mlp_set_params(dgx_mlp* m, const float* host)
{
	if (!m || !host) return DGX_ERR_ARG;
	if (hipMemcpyAsync(m->w, host, sizeof(float) * m->nparam, hipMemcpyHostToDevice, m->s) != hipSuccess) return DGX_ERR_HIP;
	if (dgx::mlp_sync_bf16(m) != hipSuccess || hipStreamSynchronize(m->s) != hipSuccess) return DGX_ERR_HIP;
	return DGX_OK;
}

dgx_status dgx_mlp_features(dgx_ctx* h, const dgx_mlp* m, const double* u, const double* inv_dt, float* feat,
                            float* sigma)
{
	if (!h || !m || !u || !inv_dt || !feat || !sigma) return DGX_ERR_ARG;
	if (!h->c.mesh_ready) return h->c.hp ? DGX_ERR_UNSUPPORTED : DGX_ERR_NO_MESH;
	if (m->nb != h->c.Np * h->c.nv || m->Np != h->c.Np) return DGX_ERR_ARG;
	return dgx::launch_features(h->c, u, inv_dt, feat, m->nf, sigma) == hipSuccess ? DGX_OK : DGX_ERR_HIP;
}

dgx_status dgx_mlp_loss_grad(dgx_mlp* m, const float* feat, const float* sigma, const double* blocks, int B,
                             double* loss, float* grad_host)
{
	if (!m || !feat || !blocks || B <= 0) return DGX_ERR_ARG;
	if (dgx::mlp_forward(m, feat, B) != hipSuccess) return DGX_ERR_HIP;
	if (dgx::mlp_loss(m, sigma, blocks, B, loss, 1.0f / B, grad_host != nullptr) != hipSuccess) return DGX_ERR_HIP;
	if (grad_host) {
		if (dgx::mlp_backward(m, feat, B) != hipSuccess) return DGX_ERR_HIP;
		if (hipMemcpyAsync(grad_host, m->gw, sizeof(float) * m->nparam, hipMemcpyDeviceToHost, m->s) != hipSuccess ||
		    hipStreamSynchronize(m->s) != hipSuccess)
			return DGX_ERR_HIP;
	}
	return DGX_OK;
}

dgx_status dgx_mlp_train(dgx_mlp* m, const float* feat, const float* sigma, const double* blocks, long long n,
                         int epochs, int batch, float lr, double* loss)
{
	if (!m || !feat || !blocks || n <= 0 || epochs < 0 || batch <= 0 || !(lr > 0.f)) return DGX_ERR_ARG;
	const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
	const long long nbatch = (n + batch - 1) / batch;
	std::vector<long long> order(nbatch);
	for (long long i = 0; i < nbatch; ++i) order[i] = i;
	uint64_t rs = 0x5EEDull + (uint64_t)m->step;
	for (int ep = 0; ep < epochs; ++ep) {
		// visit the mini-batches (contiguous runs of samples) in a fresh seeded order each epoch
		for (long long i = nbatch - 1; i > 0; --i) {
			const long long j = (long long)(dgx::splitmix64(rs) % (uint64_t)(i + 1));
			const long long t = order[i];
			order[i] = order[j];
			order[j] = t;
		}
		for (long long oi = 0; oi < nbatch; ++oi) {
			const long long bi = order[oi];
			const long long s0 = bi * batch;
			const int B = (int)((s0 + batch <= n) ? batch : n - s0);
			const float* X = feat + (size_t)s0 * m->nf;
			const float* sg = sigma ? sigma + s0 : nullptr;
			const double* A = blocks + (size_t)s0 * m->nb * m->nb;
			if (dgx::mlp_forward(m, X, B) != hipSuccess) return DGX_ERR_HIP;
			if (dgx::mlp_loss(m, sg, A, B, nullptr, 1.0f / B, true) != hipSuccess) return DGX_ERR_HIP;
			if (dgx::mlp_backward(m, X, B) != hipSuccess) return DGX_ERR_HIP;
			++m->step;
			const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
			hipLaunchKernelGGL(dgx::k_adam, dim3(dgx::gridn(m->nparam)), dim3(256), 0, m->s, m->w, m->gw, m->am, m->av,
			                   m->nparam, lr, b1, b2, c1, c2, eps);
			if (hipGetLastError() != hipSuccess || dgx::mlp_sync_bf16(m) != hipSuccess) return DGX_ERR_HIP;
		}
	}
	if (loss) {
		// full-set loss in chunks
		double tot = 0.0;
		for (long long bi = 0; bi < nbatch; ++bi) {
			const long long s0 = bi * batch;
			const int B = (int)((s0 + batch <= n) ? batch : n - s0);
			double l = 0.0;
			if (dgx::mlp_forward(m, feat + (size_t)s0 * m->nf, B) != hipSuccess ||
			    dgx::mlp_loss(m, sigma ? sigma + s0 : nullptr, blocks + (size_t)s0 * m->nb * m->nb, B, &l, 1.0f, false) !=
			        hipSuccess)
				return DGX_ERR_HIP;
			tot += l * B;
		}
		*loss = tot / n;
	}
	return hipStreamSynchronize(m->s) == hipSuccess ? DGX_OK : DGX_ERR_HIP;
}

dgx_status dgx_mlp_predict(dgx_mlp* m, const float* feat, int B, float* out)
{
	if (!m || !feat || !out || B <= 0) return DGX_ERR_ARG;
	if (dgx::mlp_forward(m, feat, B) != hipSuccess) return DGX_ERR_HIP;
	return hipMemcpyAsync(out, m->o, sizeof(float) * (size_t)B * m->nout, hipMemcpyDeviceToDevice, m->s) == hipSuccess
	           ? DGX_OK
	           : DGX_ERR_HIP;
}

dgx_status dgx_attach_mlp(dgx_ctx* h, dgx_mlp* m)
{
	if (!h) return DGX_ERR_ARG;
	h->c.mlp = m;
	return DGX_OK;
}

}  // extern "C"
