// dgx_implicit.hip -- pseudo-time Newton-GMRES for steady problems (solve_implicit.c:95-110, 232-281).
//
// Per Newton step: r = R(u); max|r| exit test (check_exit, solve_implicit.c:258-281); CFL-ramped
// per-element dt = cfl * h_min / max wave speed (compute_dt_cfl_constrained, solve_dg.c:284-319);
// linear system (dR/du - M/dt) du = -r (compute_CFL_ramping solve_dg.c:223-248, fill_petsc_Vec_b_dg
// :203-221) solved by right-preconditioned restarted GMRES (the reference uses PETSc GMRES(60),
// rtol 1e-4, external/petsc/options_files/petsc_options_gmres_default.txt) with a block-Jacobi LU or a
// learned preconditioner; update u += relax du with positivity under-relaxation (update_coef /
// compute_relaxation, solve_implicit.c:780-873).  The operator is matrix-free: exact J v by dual numbers,
// M v sum-factorised.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "dgx_device.h"
#include "dgx_internal.h"

namespace dgx {

template <int D, int N>
__device__ inline void coords2(int n, int* c)
{
	c[0] = n % N;
	c[1] = (D > 1) ? (n / N) % N : 0;
	c[2] = (D > 2) ? n / (N * N) : 0;
}
template <int D, int N>
__device__ inline double apply1d_(const double* M, const double* buf, int n, int k)
{
	int c[3];
	coords2<D, N>(n, c);
	double acc = 0.0;
	for (int j = 0; j < N; ++j) {
		int cc[3] = { c[0], c[1], c[2] };
		cc[k] = j;
		acc += M[c[k] * N + j] * buf[cc[0] + N * (cc[1] + N * cc[2])];
	}
	return acc;
}

// y_e -= s_e * M_e x_e with M_e = Ibar^T diag(w J) Ibar per variable
template <int D, int N, int NV>
__global__ void __launch_bounds__(128) k_mass_sub(const double* __restrict__ x, const double* __restrict__ detJ,
                                                  const double* __restrict__ scale, const Ops1D<double>* ops, int E,
                                                  double* y)
{
	constexpr int Np = ipow(N, D);
	__shared__ double B0[NV * Np], B1[NV * Np], sI[N * N], sIT[N * N], sW[N];
	const int e = blockIdx.x, n = threadIdx.x;
	for (int i = threadIdx.x; i < N * N; i += blockDim.x) { sI[i] = ops->I[i]; sIT[i] = ops->IT[i]; }
	for (int i = threadIdx.x; i < N; i += blockDim.x) sW[i] = ops->w[i];
	for (int i = threadIdx.x; i < NV * Np; i += blockDim.x) B0[i] = x[(size_t)e * NV * Np + i];
	__syncthreads();
	double* cur = B0;
	double* nxt = B1;
	for (int pass = 0; pass < 2 * D; ++pass) {
		const double* M = pass < D ? sI : sIT;
		if (n < Np)
			for (int v = 0; v < NV; ++v) nxt[v * Np + n] = apply1d_<D, N>(M, cur + v * Np, n, pass % D);
		__syncthreads();
		if (pass == D - 1 && n < Np) {
			int c[3];
			coords2<D, N>(n, c);
			double wq = sW[c[0]] * sW[c[1]];
			if (D == 3) wq *= sW[c[2]];
			const double s = wq * detJ[(size_t)e * Np + n];
			for (int v = 0; v < NV; ++v) nxt[v * Np + n] *= s;
		}
		__syncthreads();
		double* t = cur; cur = nxt; nxt = t;
	}
	const double s = scale ? scale[e] : 1.0;
	for (int i = threadIdx.x; i < NV * Np; i += blockDim.x) y[(size_t)e * NV * Np + i] -= s * cur[i];
}

// per-element dt = cfl * min(h_min / max(|V| + c), h_min^2 / mu) over the Bezier coefficients (the viscous
// bound with mu > 0, the Navier-Stokes constant viscosity; solve_dg.c:284-319)
template <int D, int N>
__global__ void __launch_bounds__(128) k_cfl_dt(const double* __restrict__ u, const double* __restrict__ hmin,
                                                const Ops1D<double>* ops, int E, double cfl, double mu,
                                                int sutherland, double r_s, double* inv_dt)
{
	constexpr int Np = ipow(N, D), NV = D + 2;
	__shared__ double B0[NV * Np], B1[NV * Np], sBinv[N * N], sWs[Np], sMu[Np];
	const int e = blockIdx.x, n = threadIdx.x;
	for (int i = threadIdx.x; i < N * N; i += blockDim.x) sBinv[i] = ops->Binv[i];
	for (int i = threadIdx.x; i < NV * Np; i += blockDim.x) B0[i] = u[(size_t)e * NV * Np + i];
	__syncthreads();
	double* cur = B0;
	double* nxt = B1;
	for (int k = 0; k < D; ++k) {
		if (n < Np)
			for (int v = 0; v < NV; ++v) nxt[v * Np + n] = apply1d_<D, N>(sBinv, cur + v * Np, n, k);
		__syncthreads();
		double* t = cur; cur = nxt; nxt = t;
	}
	if (n < Np) {
		const double rho = cur[n], ri = 1.0 / rho;
		double V2 = 0.0;
		for (int k = 0; k < D; ++k) { const double w = cur[(1 + k) * Np + n] * ri; V2 += w * w; }
		const double p = kGM1 * (cur[(D + 1) * Np + n] - 0.5 * rho * V2);
		sWs[n] = sqrt(V2) + sqrt(kGamma * p * ri);
		// the largest viscosity over the coefficients (compute_viscosity, solve_dg.c:300-307)
		const double T = p * ri / r_s;
		sMu[n] = sutherland ? 8e-2 * mu * sqrt(T) / (1.0 + 112.0 / T) : mu;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		double m = 0.0, mmu = 0.0;
		for (int i = 0; i < Np; ++i) {
			m = fmax(m, sWs[i]);
			mmu = fmax(mmu, sMu[i]);
		}
		inv_dt[e] = fmax(m / (cfl * hmin[e]), mmu / (cfl * hmin[e] * hmin[e]));
	}
}

// u_e += relax_e du_e, relax halved until the Bezier-coefficient means of rho and p exceed EPS_PHYS
// (compute_relaxation, solve_implicit.c:823-873)
template <int D, int N>
__global__ void __launch_bounds__(128) k_relaxed_update(double* u, const double* __restrict__ du,
                                                       const Ops1D<double>* ops, int E, int* status,
                                                       double* relax_out)
{
	constexpr int Np = ipow(N, D), NV = D + 2;
	__shared__ double Bu[NV * Np], Bd[NV * Np], T[NV * Np], sBinv[N * N];
	__shared__ double sRelax;
	const int e = blockIdx.x, n = threadIdx.x;
	for (int i = threadIdx.x; i < N * N; i += blockDim.x) sBinv[i] = ops->Binv[i];
	for (int i = threadIdx.x; i < NV * Np; i += blockDim.x) {
		Bu[i] = u[(size_t)e * NV * Np + i];
		Bd[i] = du[(size_t)e * NV * Np + i];
	}
	__syncthreads();
	for (int arr = 0; arr < 2; ++arr) {
		double* X = arr == 0 ? Bu : Bd;
		for (int k = 0; k < D; ++k) {
			if (n < Np)
				for (int v = 0; v < NV; ++v) T[v * Np + n] = apply1d_<D, N>(sBinv, X + v * Np, n, k);
			__syncthreads();
			if (n < Np)
				for (int v = 0; v < NV; ++v) X[v * Np + n] = T[v * Np + n];
			__syncthreads();
		}
	}
	if (threadIdx.x == 0) {
		double relax = 1.0;
		for (;;) {
			double ar = 0.0, ap = 0.0;
			for (int i = 0; i < Np; ++i) {
				const double rho = Bu[i] + relax * Bd[i];
				double V2 = 0.0;
				for (int k = 0; k < D; ++k) {
					const double w = (Bu[(1 + k) * Np + i] + relax * Bd[(1 + k) * Np + i]) / rho;
					V2 += w * w;
				}
				ar += rho;
				ap += kGM1 * ((Bu[(D + 1) * Np + i] + relax * Bd[(D + 1) * Np + i]) - 0.5 * rho * V2);
			}
			ar /= Np;
			ap /= Np;
			if (!(ar < kEpsPhys || ap < kEpsPhys)) break;
			relax *= 0.5;
			if (relax < kEps) { atomicOr(status, 1); break; }
		}
		sRelax = relax;
		if (relax_out) relax_out[e] = relax;
	}
	__syncthreads();
	const double r = sRelax;
	for (int i = threadIdx.x; i < NV * Np; i += blockDim.x) u[(size_t)e * NV * Np + i] += r * du[(size_t)e * NV * Np + i];
}

// ------------------------------------------------------------------------------- BLAS-1 (deterministic)
constexpr int kChunks = 64;

// partial[j*kChunks + c] = sum over chunk c of V_j . w   (j < k); V in FP64, or the FP32 Krylov basis (the
// products and sums in FP64)
template <typename TV>
__global__ void __launch_bounds__(256) k_dots_partial(const TV* __restrict__ V, size_t ldv,
                                                      const double* __restrict__ w, size_t n, double* partial)
{
	const int j = blockIdx.y, c = blockIdx.x;
	const size_t chunk = (n + kChunks - 1) / kChunks;
	const size_t lo = (size_t)c * chunk, hi = (lo + chunk < n) ? lo + chunk : n;
	const TV* vj = V + (size_t)j * ldv;
	double s = 0.0;
	for (size_t i = lo + threadIdx.x; i < hi; i += blockDim.x) s += (double)vj[i] * w[i];
	__shared__ double red[256];
	red[threadIdx.x] = s;
	__syncthreads();
	for (int off = 128; off > 0; off >>= 1) {
		if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
		__syncthreads();
	}
	if (threadIdx.x == 0) partial[(size_t)j * kChunks + c] = red[0];
}
__global__ void k_dots_final(const double* __restrict__ partial, int k, double* out)
{
	const int j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= k) return;
	double s = 0.0;
	for (int c = 0; c < kChunks; ++c) s += partial[(size_t)j * kChunks + c];
	out[j] = s;
}
// w -= V h (k vectors)
__global__ void k_gemv_sub(const double* __restrict__ V, size_t ldv, const double* __restrict__ h, int k, double* w,
                           size_t n)
{
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
		double s = 0.0;
		for (int j = 0; j < k; ++j) s += V[(size_t)j * ldv + i] * h[j];
		w[i] -= s;
	}
}
// The same reductions folded into their consumers (no k_dots_final launch): coefficient j is the sum of the
// kChunks partials of row j in chunk order, exactly as k_dots_final forms it.
__device__ inline double sum_partials(const double* __restrict__ partial, int j)
{
	double s = 0.0;
	for (int c = 0; c < kChunks; ++c) s += partial[(size_t)j * kChunks + c];
	return s;
}
// w -= V h, h reduced from the partials of V^T w by each workgroup (k <= 256)
template <typename TV>
__global__ void __launch_bounds__(256) k_gemv_sub_partials(const TV* __restrict__ V, size_t ldv,
                                                           const double* __restrict__ partial, int k, double* w, size_t n)
{
	__shared__ double h[256];
	if ((int)threadIdx.x < k) h[threadIdx.x] = sum_partials(partial, threadIdx.x);
	__syncthreads();
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
		double s = 0.0;
		for (int j = 0; j < k; ++j) s += (double)V[(size_t)j * ldv + i] * h[j];
		w[i] -= s;
	}
}
// y = x / sqrt(|x|^2), |x|^2 reduced from its partials
__global__ void k_scale_by_rnorm_partials(const double* __restrict__ x, const double* __restrict__ partial, double* y,
                                          size_t n)
{
	__shared__ double a;
	if (threadIdx.x == 0) {
		const double nv = sum_partials(partial, 0);
		a = nv > 0.0 ? 1.0 / sqrt(nv) : 0.0;
	}
	__syncthreads();
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		y[i] = a * x[i];
}
// the next Krylov basis vector: y64 = x / |x| (FP64, the preconditioner's input) and its FP32 copy y32
__global__ void k_scale_by_rnorm_partials2(const double* __restrict__ x, const double* __restrict__ partial,
                                           float* y32, double* y64, size_t n)
{
	__shared__ double a;
	if (threadIdx.x == 0) {
		const double nv = sum_partials(partial, 0);
		a = nv > 0.0 ? 1.0 / sqrt(nv) : 0.0;
	}
	__syncthreads();
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
		const double v = a * x[i];
		y64[i] = v;
		y32[i] = (float)v;
	}
}
// y = x / sqrt(nrm2[0])  (normalisation by a device-resident squared norm)
__global__ void k_scale_by_rnorm(const double* __restrict__ x, const double* __restrict__ nrm2, double* y, size_t n)
{
	const double nv = nrm2[0];
	const double a = nv > 0.0 ? 1.0 / sqrt(nv) : 0.0;
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		y[i] = a * x[i];
}
// y = a x + b y
__global__ void k_axpby(double a, const double* __restrict__ x, double b, double* y, size_t n)
{
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		y[i] = a * x[i] + (b == 0.0 ? 0.0 : b * y[i]);
}
// y = V c (k vectors), accumulate
template <typename TV>
__global__ void k_gemv_add(const TV* __restrict__ V, size_t ldv, const double* __restrict__ c, int k, double* y,
                           size_t n)
{
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
		double s = 0.0;
		for (int j = 0; j < k; ++j) s += (double)V[(size_t)j * ldv + i] * c[j];
		y[i] += s;
	}
}

static int grid_for(size_t n) { return (int)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096); }

struct Blas {
	hipStream_t s;
	double* partial = nullptr;  // kChunks * kmax
	double* partial2 = nullptr; // kChunks * kmax: second Gram-Schmidt pass
	double* partialn = nullptr; // kChunks: |w|^2
	double* dvec = nullptr;     // kmax
	std::vector<double> h;
	int kmax = 0;
	hipError_t init(hipStream_t st, int k)
	{
		s = st;
		kmax = k;
		h.resize(k);
		DGX_HIP_CHECK(hipMalloc(&partial, sizeof(double) * kChunks * k));
		DGX_HIP_CHECK(hipMalloc(&partial2, sizeof(double) * kChunks * k));
		DGX_HIP_CHECK(hipMalloc(&partialn, sizeof(double) * kChunks));
		DGX_HIP_CHECK(hipMalloc(&dvec, sizeof(double) * k));
		return hipSuccess;
	}
	void release()
	{
		for (double* q : { partial, partial2, partialn, dvec })
			if (q) (void)hipFree(q);
		partial = partial2 = partialn = dvec = nullptr;
	}
	// device out[0..k) = V^T w (deterministic two-stage reduction, no host round trip)
	hipError_t dots_dev(const double* V, size_t ldv, int k, const double* w, size_t n, double* out)
	{
		hipLaunchKernelGGL(k_dots_partial<double>, dim3(kChunks, k), dim3(256), 0, s, V, ldv, w, n, partial);
		hipLaunchKernelGGL(k_dots_final, dim3((k + 63) / 64), dim3(64), 0, s, partial, k, out);
		return hipGetLastError();
	}
	// host h[0..k) = V^T w
	hipError_t dots(const double* V, size_t ldv, int k, const double* w, size_t n, double* hout)
	{
		DGX_HIP_CHECK(dots_dev(V, ldv, k, w, n, dvec));
		DGX_HIP_CHECK(hipMemcpyAsync(hout, dvec, sizeof(double) * k, hipMemcpyDeviceToHost, s));
		return hipStreamSynchronize(s);
	}
	// partials only (rows 0..k), into buf
	template <typename TV>
	hipError_t dots_partials(const TV* V, size_t ldv, int k, const double* w, size_t n, double* buf)
	{
		hipLaunchKernelGGL(k_dots_partial<TV>, dim3(kChunks, k), dim3(256), 0, s, V, ldv, w, n, buf);
		return hipGetLastError();
	}
	// w -= V h, h = the coefficients held as partials in buf (one workgroup per CU, each reduces them once)
	template <typename TV>
	hipError_t gemv_sub_partials(const TV* V, size_t ldv, const double* buf, int k, double* w, size_t n)
	{
		const int grid = (int)((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024);
		hipLaunchKernelGGL(k_gemv_sub_partials<TV>, dim3(grid), dim3(256), 0, s, V, ldv, buf, k, w, n);
		return hipGetLastError();
	}
	hipError_t scale_by_rnorm_partials2(const double* x, const double* buf, float* y32, double* y64, size_t n)
	{
		hipLaunchKernelGGL(k_scale_by_rnorm_partials2, dim3(grid_for(n)), dim3(256), 0, s, x, buf, y32, y64, n);
		return hipGetLastError();
	}
	hipError_t scale_by_rnorm_partials(const double* x, const double* buf, double* y, size_t n)
	{
		hipLaunchKernelGGL(k_scale_by_rnorm_partials, dim3(grid_for(n)), dim3(256), 0, s, x, buf, y, n);
		return hipGetLastError();
	}
	// w -= V h with h already on the device
	hipError_t gemv_sub_dev(const double* V, size_t ldv, const double* hdev, int k, double* w, size_t n)
	{
		hipLaunchKernelGGL(k_gemv_sub, dim3(grid_for(n)), dim3(256), 0, s, V, ldv, hdev, k, w, n);
		return hipGetLastError();
	}
	hipError_t scale_by_rnorm(const double* x, const double* nrm2_dev, double* y, size_t n)
	{
		hipLaunchKernelGGL(k_scale_by_rnorm, dim3(grid_for(n)), dim3(256), 0, s, x, nrm2_dev, y, n);
		return hipGetLastError();
	}
	hipError_t gemv_sub(const double* V, size_t ldv, const double* hh, int k, double* w, size_t n)
	{
		DGX_HIP_CHECK(hipMemcpyAsync(dvec, hh, sizeof(double) * k, hipMemcpyHostToDevice, s));
		hipLaunchKernelGGL(k_gemv_sub, dim3(grid_for(n)), dim3(256), 0, s, V, ldv, dvec, k, w, n);
		return hipGetLastError();
	}
	template <typename TV>
	hipError_t gemv_add(const TV* V, size_t ldv, const double* cc, int k, double* y, size_t n)
	{
		DGX_HIP_CHECK(hipMemcpyAsync(dvec, cc, sizeof(double) * k, hipMemcpyHostToDevice, s));
		hipLaunchKernelGGL(k_gemv_add<TV>, dim3(grid_for(n)), dim3(256), 0, s, V, ldv, dvec, k, y, n);
		return hipGetLastError();
	}
	hipError_t axpby(double a, const double* x, double b, double* y, size_t n)
	{
		hipLaunchKernelGGL(k_axpby, dim3(grid_for(n)), dim3(256), 0, s, a, x, b, y, n);
		return hipGetLastError();
	}
	hipError_t norm(const double* x, size_t n, double* out)
	{
		double v;
		DGX_HIP_CHECK(dots(x, 0, 1, x, n, &v));
		*out = std::sqrt(v);
		return hipSuccess;
	}
};

// one workgroup per element, a lane per node (HEX p = 4: 128 lanes for its 125 nodes)
#define DGX_PER_ELEM2(KNAME, BLK, ...)                                                               \
	switch (c.d * 16 + c.N) {                                                                       \
	case 2 * 16 + 2: hipLaunchKernelGGL((KNAME<2, 2 __VA_ARGS__>), dim3(c.E), dim3(BLK), 0, c.stream, ARGS); break; \
	case 2 * 16 + 3: hipLaunchKernelGGL((KNAME<2, 3 __VA_ARGS__>), dim3(c.E), dim3(BLK), 0, c.stream, ARGS); break; \
	case 2 * 16 + 4: hipLaunchKernelGGL((KNAME<2, 4 __VA_ARGS__>), dim3(c.E), dim3(BLK), 0, c.stream, ARGS); break; \
	case 2 * 16 + 5: hipLaunchKernelGGL((KNAME<2, 5 __VA_ARGS__>), dim3(c.E), dim3(BLK), 0, c.stream, ARGS); break; \
	case 2 * 16 + 6: hipLaunchKernelGGL((KNAME<2, 6 __VA_ARGS__>), dim3(c.E), dim3(BLK), 0, c.stream, ARGS); break; \
	case 3 * 16 + 2: hipLaunchKernelGGL((KNAME<3, 2 __VA_ARGS__>), dim3(c.E), dim3(64), 0, c.stream, ARGS); break;  \
	case 3 * 16 + 3: hipLaunchKernelGGL((KNAME<3, 3 __VA_ARGS__>), dim3(c.E), dim3(64), 0, c.stream, ARGS); break;  \
	case 3 * 16 + 4: hipLaunchKernelGGL((KNAME<3, 4 __VA_ARGS__>), dim3(c.E), dim3(64), 0, c.stream, ARGS); break;  \
	case 3 * 16 + 5: hipLaunchKernelGGL((KNAME<3, 5 __VA_ARGS__>), dim3(c.E), dim3(128), 0, c.stream, ARGS); break; \
	default: return hipErrorInvalidValue;                                                       \
	}

hipError_t launch_mass_sub(Ctx& c, const double* x, const double* scale, double* y)
{
	if (c.E == 0) return hipSuccess;
	if (c.d != 2 && c.d != 3) return hipErrorInvalidValue;
#define ARGS x, c.detJ_vc, scale, c.ops64, c.E, y
	if (c.nv == 1) { DGX_PER_ELEM2(k_mass_sub, 64, , 1) }
	else if (c.nv == 4) { DGX_PER_ELEM2(k_mass_sub, 64, , 4) }
	else { DGX_PER_ELEM2(k_mass_sub, 64, , 5) }
#undef ARGS
	return hipGetLastError();
}

hipError_t launch_cfl_dt(Ctx& c, const double* u, double cfl, double* inv_dt)
{
	if (c.nv != c.d + 2) return hipErrorInvalidValue;   // Euler and Navier-Stokes
#define ARGS u, c.hmin, c.ops64, c.E, cfl, (c.cfg.pde == DGX_PDE_NAVIER_STOKES ? c.cfg.mu : 0.0),                   \
             (c.cfg.pde == DGX_PDE_NAVIER_STOKES && c.cfg.viscosity_type == DGX_VISC_SUTHERLAND) ? 1 : 0,          \
             (c.cfg.r_s > 0.0 ? c.cfg.r_s : 1.0), inv_dt
	DGX_PER_ELEM2(k_cfl_dt, 64, )
#undef ARGS
	return hipGetLastError();
}

hipError_t launch_relaxed_update(Ctx& c, double* u, const double* du, int* status)
{
	if (c.nv != c.d + 2) return hipErrorInvalidValue;
#define ARGS u, du, c.ops64, c.E, status, (double*)nullptr
	DGX_PER_ELEM2(k_relaxed_update, 64, )
#undef ARGS
	return hipGetLastError();
}

}  // namespace dgx

// ------------------------------------------------------------------------------- preconditioners
namespace dgx {

// 2-D face-node map: node of local face lf at face position a (faces r-, r+, s-, s+)
template <int D>
__device__ inline int sgs_face_node(int lf, int a, int N)
{
	const int fd = lf >> 1, fix = (lf & 1) ? N - 1 : 0;
	if constexpr (D == 2) return fd == 0 ? fix + N * a : a + N * fix;
	const int i = a % N, j = a / N;   // HEX face position a = i + N j over the two remaining directions
	if (fd == 0) return fix + N * (i + N * j);
	if (fd == 1) return i + N * (fix + N * j);
	return i + N * (j + N * fix);
}

// One colour of a block Gauss-Seidel sweep: for every element e of the colour,
// z_e = A_ee^-1 (v_e - sum_faces A_e,n z_n).  A_ee^-1 is stored transposed (nb x nb); the face blocks A_e,n
// compactly, transposed, on the nfv = nv N^(D-1) face-node rows and columns (the MODE_OFFD layout), so a face
// costs an nfv x nfv mat-vec on the neighbour's trace.  Elements of one colour share no face, so their
// updates are independent.  One workgroup of 64 (QUAD) or 128 (HEX) lanes per element, rows and face dofs beyond
// the lanes looping; nb <= kMaxInvBlock and 2 D nfv <= kMaxSgsFace.
// TB: the stored blocks' type (FP64, or FP32 with the products and sums in FP64).
template <typename TB, int D>
__global__ void __launch_bounds__(D == 2 ? 64 : 128) k_bgs_color(const int* __restrict__ elems,
                                                               const ElemFace* __restrict__ ef,
                                                               const TB* __restrict__ invT, const TB* __restrict__ offT,
                                                               const double* __restrict__ rhs, double* z, int N, int nv)
{
	constexpr int NFACE = 2 * D, NT = (D == 2) ? 64 : 128;
	__shared__ double sz[kMaxSgsFace], sf[kMaxSgsFace], sr[kMaxInvBlock];
	const int NF = (D == 2) ? N : N * N, Np = NF * N, nb = Np * nv, nfv = nv * NF;
	const int e = elems[blockIdx.x], t = threadIdx.x;
	for (int it = t; it < NFACE * nfv; it += NT) {   // neighbour traces, in each neighbour's own face-node order
		const int lf = it / nfv, c = it - lf * nfv, w = c / NF;
		const ElemFace f = ef[(size_t)e * NFACE + lf];
		sz[it] = (f.en >= 0) ? z[(size_t)f.en * nb + w * Np + sgs_face_node<D>(f.packed & 15, c - w * NF, N)] : 0.0;
	}
	__syncthreads();
	for (int it = t; it < NFACE * nfv; it += NT) {   // face contributions on the element's face-node rows
		const int lf = it / nfv, row = it - lf * nfv;
		const TB* C = offT + ((size_t)e * NFACE + lf) * nfv * nfv;
		double acc = 0.0;
		for (int c = 0; c < nfv; ++c) acc += (double)C[(size_t)c * nfv + row] * sz[lf * nfv + c];
		sf[it] = acc;
	}
	__syncthreads();
	for (int i = t; i < nb; i += NT) {
		const int v = i / Np, node = i - v * Np, i0 = node % N;
		double r = rhs[(size_t)e * nb + i];
		if constexpr (D == 2) {
			const int i1 = node / N;
			if (i0 == 0) r -= sf[0 * nfv + v * N + i1];
			if (i0 == N - 1) r -= sf[1 * nfv + v * N + i1];
			if (i1 == 0) r -= sf[2 * nfv + v * N + i0];
			if (i1 == N - 1) r -= sf[3 * nfv + v * N + i0];
		} else {
			const int i1 = (node / N) % N, i2 = node / (N * N);
			if (i0 == 0) r -= sf[0 * nfv + v * NF + i1 + N * i2];
			if (i0 == N - 1) r -= sf[1 * nfv + v * NF + i1 + N * i2];
			if (i1 == 0) r -= sf[2 * nfv + v * NF + i0 + N * i2];
			if (i1 == N - 1) r -= sf[3 * nfv + v * NF + i0 + N * i2];
			if (i2 == 0) r -= sf[4 * nfv + v * NF + i0 + N * i1];
			if (i2 == N - 1) r -= sf[5 * nfv + v * NF + i0 + N * i1];
		}
		sr[i] = r;
	}
	__syncthreads();
	for (int i = t; i < nb; i += NT) {
		const TB* M = invT + (size_t)e * nb * nb;
		double acc = 0.0;
		for (int k = 0; k < nb; ++k) acc += (double)M[(size_t)k * nb + i] * sr[k];
		z[(size_t)e * nb + i] = acc;
	}
}

// The same colour sweep with full face blocks (the BR2 residual: the neighbour's whole state enters through its
// lifted gradient): offT holds dR_e/du_n transposed per local face, [e][lf][col][row], nb <= kMaxInvBlock.  Four
// wavefronts per element, one per local face for the face products and one quarter of the columns each for the
// inverse, so that four times as many loads are in flight as with one wavefront (the sweep reads 4 nb^2 + nb^2
// values per element and is bound by load latency, not bytes); rows beyond 64 loop.
template <typename TB>
__global__ void __launch_bounds__(256) k_bgs_full_color(const int* __restrict__ elems, const ElemFace* __restrict__ ef,
                                                        const TB* __restrict__ invT, const TB* __restrict__ offT,
                                                        const double* __restrict__ rhs, double* z, int nb)
{
	__shared__ double sz[4][kMaxInvBlock], sf[4][kMaxInvBlock], sr[kMaxInvBlock];
	const int e = elems[blockIdx.x], t = threadIdx.x, lf = t >> 6, l = t & 63;
	{
		const ElemFace f = ef[(size_t)e * 4 + lf];
		for (int row = l; row < nb; row += 64) sz[lf][row] = (f.en >= 0) ? z[(size_t)f.en * nb + row] : 0.0;
	}
	__syncthreads();
	for (int row = l; row < nb; row += 64) {   // face lf's product, row `row`
		const TB* C = offT + ((size_t)e * 4 + lf) * nb * nb;
		double acc = 0.0;
		for (int c = 0; c < nb; ++c) acc += (double)C[(size_t)c * nb + row] * sz[lf][c];
		sf[lf][row] = acc;
	}
	__syncthreads();
	for (int i = t; i < nb; i += 256) sr[i] = rhs[(size_t)e * nb + i] - ((sf[0][i] + sf[1][i]) + (sf[2][i] + sf[3][i]));
	__syncthreads();
	for (int row = l; row < nb; row += 64) {   // quarter lf of the columns of A_ee^-1 r
		const TB* M = invT + (size_t)e * nb * nb;
		const int k0 = lf * ((nb + 3) / 4), k1 = min(nb, k0 + (nb + 3) / 4);
		double acc = 0.0;
		for (int k = k0; k < k1; ++k) acc += (double)M[(size_t)k * nb + row] * sr[k];
		sf[lf][row] = acc;
	}
	__syncthreads();
	for (int i = t; i < nb; i += 256) z[(size_t)e * nb + i] = (sf[0][i] + sf[1][i]) + (sf[2][i] + sf[3][i]);
}

struct PrecondIdentity : Precond {
	hipError_t setup(Ctx&, const double*, const double*) override { return hipSuccess; }
	hipError_t apply(Ctx& c, const double* v, double* z) override
	{
		return hipMemcpyAsync(z, v, sizeof(double) * (size_t)c.E * c.Np * c.nv, hipMemcpyDeviceToDevice, c.stream);
	}
};

// exact inverse (or LU) of the element diagonal blocks of (dR/du - M/dt): the block-Jacobi baseline
hipError_t launch_features(Ctx& c, const double* u, const double* inv_dt, float* feat, int nf, float* sigma);

struct PrecondBlockJacobi : Precond {
	double* blocks = nullptr;
	int* piv = nullptr;
	// optional training-set capture
	double* cap_blocks = nullptr;
	float* cap_feat = nullptr;
	float* cap_sigma = nullptr;
	long long cap_capacity = 0, cap_count = 0;
	~PrecondBlockJacobi() override
	{
		if (blocks) (void)hipFree(blocks);
		if (piv) (void)hipFree(piv);
	}
	hipError_t setup(Ctx& c, const double* u, const double* inv_dt) override
	{
		const int nb = c.Np * c.nv;
		if (!blocks) {
			DGX_HIP_CHECK(hipMalloc(&blocks, sizeof(double) * (size_t)c.E * nb * nb));
			DGX_HIP_CHECK(hipMalloc(&piv, sizeof(int) * (size_t)c.E * nb));
		}
		for (int col = 0; col < nb; ++col) DGX_HIP_CHECK(launch_diag_col(c, u, blocks, col));
		DGX_HIP_CHECK(launch_sub_mass(c, blocks, inv_dt, 0.0));
		if (cap_blocks && cap_feat && cap_sigma && cap_count + c.E <= cap_capacity) {
			const int nf = (nb + c.Np + 1 + 15) / 16 * 16;
			DGX_HIP_CHECK(hipMemcpyAsync(cap_blocks + (size_t)cap_count * nb * nb, blocks,
			                             sizeof(double) * (size_t)c.E * nb * nb, hipMemcpyDeviceToDevice, c.stream));
			DGX_HIP_CHECK(launch_features(c, u, inv_dt, cap_feat + (size_t)cap_count * nf, nf,
			                              cap_sigma + (size_t)cap_count));
			cap_count += c.E;
		}
		// blocks up to 142 x 142: explicit inverse (one coalesced mat-vec per application); larger: LU
		if (nb <= kMaxInvBlock) return launch_block_inverse(c.stream, blocks, c.E, nb, nullptr);
		return launch_block_lu(c.stream, blocks, piv, c.E, nb);
	}
	hipError_t apply(Ctx& c, const double* v, double* z) override
	{
		const int nb = c.Np * c.nv;
		if (nb <= kMaxInvBlock) return launch_block_inv_apply(c.stream, blocks, v, z, c.E, nb);
		DGX_HIP_CHECK(hipMemcpyAsync(z, v, sizeof(double) * (size_t)c.E * nb, hipMemcpyDeviceToDevice, c.stream));
		return launch_block_lu_solve(c.stream, blocks, piv, z, c.E, nb);
	}
};

// Multicolour block symmetric Gauss-Seidel on the block-sparse Jacobian (diagonal blocks with the -M/dt
// term, exact inverses; compact face blocks from the dual-number MODE_OFFD columns): forward over the
// colours, then backward, from z = 0.  A fixed linear operator, so plain right-preconditioned GMRES applies.
struct PrecondBlockSGS : Precond {
	PrecondBlockJacobi diag;   // assembles and inverts A_ee - M_e/dt
	double* offT = nullptr;    // E * 4 * nfv^2 compact transposed face blocks (Navier-Stokes: E * 4 * nb^2, full)
	bool fp32 = false;         // sweep over FP32 copies of the inverse and face blocks
	float* inv32 = nullptr;
	float* off32 = nullptr;
	explicit PrecondBlockSGS(bool use_fp32 = false) : fp32(use_fp32) {}
	~PrecondBlockSGS() override
	{
		if (offT) (void)hipFree(offT);
		if (inv32) (void)hipFree(inv32);
		if (off32) (void)hipFree(off32);
	}
	static bool full(const Ctx& c) { return c.cfg.pde == DGX_PDE_NAVIER_STOKES; }
	hipError_t setup(Ctx& c, const double* u, const double* inv_dt) override
	{
		const int nb = c.Np * c.nv, nf = 2 * c.d, nfv = c.nv * (c.d == 2 ? c.N : c.N * c.N);
		if ((c.d != 2 && c.d != 3) || (full(c) && c.d != 2) || nb > kMaxInvBlock || nf * nfv > kMaxSgsFace)
			return hipErrorInvalidValue;
		DGX_HIP_CHECK(diag.setup(c, u, inv_dt));
		// the face blocks: compact on the face nodes (Euler, advection), full for BR2
		const int fb = full(c) ? nb : nfv;
		if (!offT) DGX_HIP_CHECK(hipMalloc(&offT, sizeof(double) * (size_t)c.E * nf * fb * fb));
		if (full(c))
			for (int col = 0; col < nb; ++col) DGX_HIP_CHECK(launch_ns_offd_col(c, u, offT, col, true));
		else
			for (int col = 0; col < nfv; ++col) DGX_HIP_CHECK(launch_offd_col(c, u, offT, col));
		if (!fp32) return hipSuccess;
		const size_t ni = (size_t)c.E * nb * nb, no = (size_t)c.E * nf * fb * fb;
		if (!inv32) DGX_HIP_CHECK(hipMalloc(&inv32, sizeof(float) * ni));
		if (!off32) DGX_HIP_CHECK(hipMalloc(&off32, sizeof(float) * no));
		DGX_HIP_CHECK(launch_convert(diag.blocks, inv32, ni, c.stream));
		return launch_convert(offT, off32, no, c.stream);
	}
	hipError_t sweep(Ctx& c, const double* v, double* z, int k)
	{
		const int n0 = c.color_off[k], cnt = c.color_off[k + 1] - n0;
		if (cnt == 0) return hipSuccess;
		if (full(c)) {
			const int nb = c.Np * c.nv;
			if (fp32)
				hipLaunchKernelGGL(k_bgs_full_color<float>, dim3(cnt), dim3(256), 0, c.stream, c.color_elems + n0,
				                   c.elem_faces, inv32, off32, v, z, nb);
			else
				hipLaunchKernelGGL(k_bgs_full_color<double>, dim3(cnt), dim3(256), 0, c.stream, c.color_elems + n0,
				                   c.elem_faces, diag.blocks, offT, v, z, nb);
			return hipGetLastError();
		}
		if (c.d == 3) {
			if (fp32)
				hipLaunchKernelGGL((k_bgs_color<float, 3>), dim3(cnt), dim3(128), 0, c.stream, c.color_elems + n0,
				                   c.elem_faces, inv32, off32, v, z, c.N, c.nv);
			else
				hipLaunchKernelGGL((k_bgs_color<double, 3>), dim3(cnt), dim3(128), 0, c.stream, c.color_elems + n0,
				                   c.elem_faces, diag.blocks, offT, v, z, c.N, c.nv);
		} else if (fp32)
			hipLaunchKernelGGL((k_bgs_color<float, 2>), dim3(cnt), dim3(64), 0, c.stream, c.color_elems + n0, c.elem_faces,
			                   inv32, off32, v, z, c.N, c.nv);
		else
			hipLaunchKernelGGL((k_bgs_color<double, 2>), dim3(cnt), dim3(64), 0, c.stream, c.color_elems + n0, c.elem_faces,
			                   diag.blocks, offT, v, z, c.N, c.nv);
		return hipGetLastError();
	}
	hipError_t apply(Ctx& c, const double* v, double* z) override
	{
		const size_t n = (size_t)c.E * c.Np * c.nv;
		DGX_HIP_CHECK(launch_fill_u64(c.stream, reinterpret_cast<unsigned long long*>(z), (int)n, 0ull));
		const int nc = (int)c.color_off.size() - 1;
		for (int k = 0; k < nc; ++k) DGX_HIP_CHECK(sweep(c, v, z, k));
		// backward: the last colour's inputs have not changed since its forward update, so it is not redone
		for (int k = nc - 2; k >= 0; --k) DGX_HIP_CHECK(sweep(c, v, z, k));
		return hipSuccess;
	}
};


// ------------------------------------------------------------------------------- GMRES
constexpr int kMaxRestart = 200;

// One Arnoldi column on the device: H(0..j, j) = h1 + h2 (the two Gram-Schmidt passes), H(j+1, j) = |w|, the
// previous Givens rotations applied to the new column, the new rotation, and the residual estimate.
// st[0] = |g(j+1)|, st[1] = first column count at which |g| <= tol (or -1), st[2] = |w|.
// The rotation chain is sequential: the column and the rotations are staged in LDS by all lanes first, so the
// chain runs on LDS values instead of waiting on global loads at every step.
__global__ void __launch_bounds__(64) k_arnoldi_step(double* H, double* cs, double* sn, double* g,
                                                    const double* __restrict__ p1, const double* __restrict__ p2,
                                                    const double* __restrict__ pn, int j, int m, double tol, double* st)
{
	__shared__ double scol[kMaxRestart + 1], sc[kMaxRestart], ss[kMaxRestart];
	const int t = threadIdx.x;
	// Gram-Schmidt coefficients of both passes and |w|^2, reduced from their partials
	for (int i = t; i <= j; i += blockDim.x) scol[i] = sum_partials(p1, i) + sum_partials(p2, i);
	for (int i = t; i < j; i += blockDim.x) {
		sc[i] = cs[i];
		ss[i] = sn[i];
	}
	__syncthreads();
	if (t == 0) {
		const double hn = sqrt(sum_partials(pn, 0));
		double a = scol[0];
		for (int i = 0; i < j; ++i) {
			const double b = scol[i + 1];
			scol[i] = sc[i] * a + ss[i] * b;
			a = -ss[i] * a + sc[i] * b;
		}
		const double b = hn;
		const double r = sqrt(a * a + b * b);
		const double c = r > 0 ? a / r : 1.0, sj = r > 0 ? b / r : 0.0;
		cs[j] = c;
		sn[j] = sj;
		scol[j] = r;
		scol[j + 1] = 0.0;
		const double gj = g[j];
		g[j + 1] = -sj * gj;
		g[j] = c * gj;
		st[0] = fabs(g[j + 1]);
		if (st[1] < 0.0 && (st[0] <= tol || hn == 0.0)) st[1] = (double)(j + 1);
		st[2] = hn;
	}
	__syncthreads();
	double* col = H + (size_t)j * (m + 1);   // column j, rows 0..m
	for (int i = t; i <= j + 1; i += blockDim.x) col[i] = scol[i];
}

// g = (beta, 0, ...), st = (beta, -1, 1) at the start of a restart cycle
__global__ void k_arnoldi_start(double* g, int m, double beta, double* st)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i <= m) g[i] = (i == 0) ? beta : 0.0;
	if (i == 0) { st[0] = beta; st[1] = -1.0; st[2] = 1.0; }
}

struct Gmres {
	int m = 60;
	static constexpr int kCheck = 4;   // Arnoldi steps between host convergence checks
	// Krylov basis (m+1) x n in FP32: the Gram-Schmidt passes read it twice per iteration, and FP32 halves
	// those bytes; every product and sum stays in FP64, and so do w, x and the newest basis vector vcur (the
	// preconditioner's input).  The Arnoldi relation then holds to FP32 rounding (~1e-7 relative), far below the
	// inner tolerance; the outer Newton residual decides convergence in FP64.
	float* V = nullptr;
	double* vcur = nullptr;
	double* w = nullptr;
	double* z = nullptr;
	double* x = nullptr;
	size_t n = 0;
	Blas blas;
	std::vector<double> H, g, cs, sn, y;
	double* dls = nullptr;    // device: H ((m+1) x m, column-major) | cs (m) | sn (m) | g (m+1) | st (3)
	hipError_t init(Ctx& c, int restart)
	{
		m = restart;
		n = (size_t)c.E * c.Np * c.nv;
		DGX_HIP_CHECK(hipMalloc(&V, sizeof(float) * (size_t)(m + 1) * n));
		DGX_HIP_CHECK(hipMalloc(&vcur, sizeof(double) * n));
		DGX_HIP_CHECK(hipMalloc(&w, sizeof(double) * n));
		DGX_HIP_CHECK(hipMalloc(&z, sizeof(double) * n));
		DGX_HIP_CHECK(hipMalloc(&x, sizeof(double) * n));
		H.assign((size_t)(m + 1) * m, 0.0);
		g.assign(m + 1, 0.0);
		cs.assign(m, 0.0);
		sn.assign(m, 0.0);
		y.assign(m, 0.0);
		DGX_HIP_CHECK(hipMalloc(&dls, sizeof(double) * ls_size()));
		return blas.init(c.stream, m + 1);
	}
	void release()
	{
		if (V) (void)hipFree(V);
		V = nullptr;
		for (double* p : { vcur, w, z, x, dls })
			if (p) (void)hipFree(p);
		vcur = w = z = x = dls = nullptr;
		blas.release();
	}
	size_t ls_size() const { return (size_t)(m + 1) * m + 2 * (size_t)m + (m + 1) + 3; }
	bool jvp32 = false;   // the operator in FP32 dual numbers
	// w = A v with A = dR/du(u) - M/dt
	hipError_t op(Ctx& c, const double* u, const double* inv_dt, const double* v, double* out)
	{
		// dual-number J v with the -M/dt term added in the same kernel (at the GL nodes, before Ibar^T)
		return jvp32 ? launch_jvp32(c, u, v, out, inv_dt) : launch_jvp(c, u, v, out, inv_dt);
	}
	// right-preconditioned restarted GMRES on A x = b, x0 = 0; returns iterations, sets converged.
	// The Hessenberg column and its Givens rotation are formed on the device; the host reads the residual
	// estimate every kCheck steps (steps past convergence only enlarge the Krylov space) and brings H and g
	// back once per restart cycle for the triangular solve.
	hipError_t solve(Ctx& c, const double* u, const double* inv_dt, const double* b, Precond& P, double rtol,
	                 int max_it, int* iters, bool* converged)
	{
		DGX_HIP_CHECK(blas.axpby(0.0, b, 0.0, x, n));
		double bnorm;
		DGX_HIP_CHECK(blas.norm(b, n, &bnorm));
		*iters = 0;
		*converged = (bnorm == 0.0);
		if (*converged) return hipSuccess;
		const double tol = rtol * bnorm;
		double* dH = dls;
		double* dcs = dH + (size_t)(m + 1) * m;
		double* dsn = dcs + m;
		double* dg = dsn + m;
		double* dst = dg + (m + 1);
		std::vector<double> hls(ls_size());
		int total = 0;
		while (total < max_it) {
			// r = b - A x  (x = 0 on the first cycle)
			if (total == 0) {
				DGX_HIP_CHECK(hipMemcpyAsync(vcur, b, sizeof(double) * n, hipMemcpyDeviceToDevice, c.stream));
			} else {
				DGX_HIP_CHECK(op(c, u, inv_dt, x, w));
				DGX_HIP_CHECK(blas.axpby(1.0, b, 0.0, vcur, n));
				DGX_HIP_CHECK(blas.axpby(-1.0, w, 1.0, vcur, n));
			}
			double beta;
			DGX_HIP_CHECK(blas.norm(vcur, n, &beta));
			if (beta <= tol) { *converged = true; break; }
			DGX_HIP_CHECK(blas.axpby(1.0 / beta, vcur, 0.0, vcur, n));
			DGX_HIP_CHECK(launch_convert(vcur, V, n, c.stream));
			hipLaunchKernelGGL(k_arnoldi_start, dim3((m + 64) / 64), dim3(64), 0, c.stream, dg, m, beta, dst);
			DGX_HIP_CHECK(hipGetLastError());
			int j = 0, jconv = -1;
			for (; j < m && total < max_it; ++j, ++total) {
				DGX_HIP_CHECK(P.apply(c, vcur, z));   // vcur = v_j in FP64
				DGX_HIP_CHECK(op(c, u, inv_dt, z, w));
				// classical Gram-Schmidt, twice (CGS2); the coefficients stay as partial sums on the device and
				// each consumer reduces them
				DGX_HIP_CHECK(blas.dots_partials(V, n, j + 1, w, n, blas.partial));
				DGX_HIP_CHECK(blas.gemv_sub_partials(V, n, blas.partial, j + 1, w, n));
				DGX_HIP_CHECK(blas.dots_partials(V, n, j + 1, w, n, blas.partial2));
				DGX_HIP_CHECK(blas.gemv_sub_partials(V, n, blas.partial2, j + 1, w, n));
				DGX_HIP_CHECK(blas.dots_partials(w, 0, 1, w, n, blas.partialn));
				DGX_HIP_CHECK(blas.scale_by_rnorm_partials2(w, blas.partialn, V + (size_t)(j + 1) * n, vcur, n));
				hipLaunchKernelGGL(k_arnoldi_step, dim3(1), dim3(64), 0, c.stream, dH, dcs, dsn, dg, blas.partial,
				                   blas.partial2, blas.partialn, j, m, tol, dst);
				DGX_HIP_CHECK(hipGetLastError());
				if ((j + 1) % kCheck == 0 || j + 1 == m || total + 1 >= max_it) {
					double st[3];
					DGX_HIP_CHECK(hipMemcpyAsync(st, dst, sizeof(st), hipMemcpyDeviceToHost, c.stream));
					DGX_HIP_CHECK(hipStreamSynchronize(c.stream));
					if (st[1] >= 0.0) { jconv = (int)st[1]; ++j; ++total; *converged = true; break; }
				}
			}
			// y = H^-1 g over the first jj columns, x += P^-1 (V y)
			const int jj = (jconv > 0) ? jconv : j;
			DGX_HIP_CHECK(hipMemcpyAsync(hls.data(), dls, sizeof(double) * hls.size(), hipMemcpyDeviceToHost, c.stream));
			DGX_HIP_CHECK(hipStreamSynchronize(c.stream));
			const double* Hh = hls.data();
			const double* gh = Hh + (size_t)(m + 1) * m + 2 * (size_t)m;
			for (int i = jj - 1; i >= 0; --i) {
				double sacc = gh[i];
				for (int k = i + 1; k < jj; ++k) sacc -= Hh[(size_t)k * (m + 1) + i] * y[k];
				y[i] = sacc / Hh[(size_t)i * (m + 1) + i];
			}
			DGX_HIP_CHECK(blas.axpby(0.0, b, 0.0, w, n));
			DGX_HIP_CHECK(blas.gemv_add(V, n, y.data(), jj, w, n));
			DGX_HIP_CHECK(P.apply(c, w, z));
			DGX_HIP_CHECK(blas.axpby(1.0, z, 1.0, x, n));
			if (*converged) break;
		}
		*iters = total;
		return hipSuccess;
	}
};

}  // namespace dgx

using dgx::Ctx;

// 3-D (HEX Euler): the FP64 J v, and no preconditioner, block-Jacobi with blocks the batched LU takes
// (nb = 5 (p + 1)^3 <= 143: p <= 2), or block SGS with blocks the in-LDS inverse and the sweep take (p <= 2)
static bool implicit_3d_ok(const Ctx& c, int precond, int jvp_fp32)
{
	if (c.d != 3) return true;
	if (jvp_fp32) return false;
	if (precond == DGX_PC_NONE) return true;
	if (precond == DGX_PC_BLOCK_SGS || precond == DGX_PC_BLOCK_SGS32)   // in-LDS inverses (nb <= 142: p <= 2)
		return c.Np * c.nv <= dgx::kMaxInvBlock && 6 * c.nv * c.N * c.N <= dgx::kMaxSgsFace;
	return precond == DGX_PC_BLOCK_JACOBI && c.Np * c.nv <= dgx::kMaxLuBlock;
}

extern "C" dgx_status dgx_precond_apply(dgx_ctx* h, int precond, const double* u, const double* inv_dt,
                                        const double* v, double* z)
{
	if (!h || !u || !inv_dt || !v || !z || v == z) return DGX_ERR_ARG;
	Ctx& c = h->c;
	if (!c.mesh_ready) return c.hp ? DGX_ERR_UNSUPPORTED : DGX_ERR_NO_MESH;
	if ((c.d != 2 && c.d != 3) || c.cfg.pde == DGX_PDE_ADVECTION) return DGX_ERR_UNSUPPORTED;
	// Navier-Stokes: no learned preconditioner (it is trained on Euler blocks)
	if (c.cfg.pde == DGX_PDE_NAVIER_STOKES && precond == DGX_PC_LEARNED) return DGX_ERR_UNSUPPORTED;
	if (!implicit_3d_ok(c, precond, 0)) return DGX_ERR_UNSUPPORTED;
	dgx::Precond* P = nullptr;
	if (precond == DGX_PC_NONE) P = new dgx::PrecondIdentity();
	else if (precond == DGX_PC_BLOCK_JACOBI) P = new dgx::PrecondBlockJacobi();
	else if (precond == DGX_PC_BLOCK_SGS) P = new dgx::PrecondBlockSGS();
	else if (precond == DGX_PC_BLOCK_SGS32) P = new dgx::PrecondBlockSGS(true);
	else if (precond == DGX_PC_LEARNED) P = dgx::make_learned_precond(c);
	if (!P) return precond == DGX_PC_LEARNED ? DGX_ERR_UNSUPPORTED : DGX_ERR_ARG;
	hipError_t e = P->setup(c, u, inv_dt);
	if (e == hipSuccess) e = P->apply(c, v, z);
	if (e == hipSuccess) e = hipStreamSynchronize(c.stream);   // the preconditioner's buffers go with it
	delete P;
	if (e == hipErrorInvalidValue) return DGX_ERR_UNSUPPORTED;
	return e == hipSuccess ? DGX_OK : DGX_ERR_HIP;
}

extern "C" dgx_status dgx_solve_implicit(dgx_ctx* h, double* u, const dgx_implicit_params* prm, dgx_implicit_stats* st)
{
	if (!h || !u || !prm || !st) return DGX_ERR_ARG;
	Ctx& c = h->c;
	if (!c.mesh_ready) return c.hp ? DGX_ERR_UNSUPPORTED : DGX_ERR_NO_MESH;
	if ((c.d != 2 && c.d != 3) || c.cfg.pde == DGX_PDE_ADVECTION) return DGX_ERR_UNSUPPORTED;
	// Navier-Stokes: no learned preconditioner (it is trained on Euler blocks)
	if (c.cfg.pde == DGX_PDE_NAVIER_STOKES && prm->precond == DGX_PC_LEARNED) return DGX_ERR_UNSUPPORTED;
	if (!implicit_3d_ok(c, prm->precond, prm->jvp_fp32)) return DGX_ERR_UNSUPPORTED;
	if (prm->gmres_restart < 1 || prm->gmres_restart > 200 || prm->max_newton < 1 || !(prm->cfl_initial > 0.0))
		return DGX_ERR_ARG;
	std::memset(st, 0, sizeof(*st));
	const auto t0 = std::chrono::steady_clock::now();
	const size_t n = (size_t)c.E * c.Np * c.nv;
	double *r = nullptr, *inv_dt = nullptr;
	int* status = nullptr;
	dgx::Gmres gm;
	dgx::Precond* P = nullptr;
	dgx::PrecondBlockJacobi* Pfb = nullptr;   // exact fallback of the learned preconditioner
	dgx_status ret = DGX_OK;
	auto fail = [&](dgx_status s) { ret = s; };
	do {
		if (hipMalloc(&r, sizeof(double) * n) != hipSuccess || hipMalloc(&inv_dt, sizeof(double) * c.E) != hipSuccess ||
		    hipMalloc(&status, sizeof(int)) != hipSuccess || gm.init(c, prm->gmres_restart) != hipSuccess) {
			fail(DGX_ERR_HIP);
			break;
		}
		gm.jvp32 = prm->jvp_fp32 != 0;
		if (prm->precond == DGX_PC_BLOCK_JACOBI) {
			auto* bj = new dgx::PrecondBlockJacobi();
			bj->cap_blocks = prm->capture_blocks;
			bj->cap_feat = prm->capture_features;
			bj->cap_sigma = prm->capture_sigma;
			bj->cap_capacity = prm->capture_capacity;
			P = bj;
		}
		else if (prm->precond == DGX_PC_BLOCK_SGS) P = new dgx::PrecondBlockSGS();
		else if (prm->precond == DGX_PC_BLOCK_SGS32) P = new dgx::PrecondBlockSGS(true);
		else if (prm->precond == DGX_PC_LEARNED) {
			P = dgx::make_learned_precond(c);
			if (prm->learned_fallback_iters > 0) Pfb = new dgx::PrecondBlockJacobi();
		}
		else P = new dgx::PrecondIdentity();
		if (!P) { fail(DGX_ERR_UNSUPPORTED); break; }
		double max_rhs0 = 0.0;
		for (int it = 0; it < prm->max_newton; ++it) {
			if (dgx::launch_rhs64(c, u, r, dgx::MODE_RHS, nullptr) != hipSuccess) { fail(DGX_ERR_HIP); break; }
			double mx = 0.0;
			if (dgx::launch_fill_u64(c.stream, c.dmax, 1, 0ull) != hipSuccess ||
			    dgx::launch_maxabs(c, r, n, c.dmax) != hipSuccess) { fail(DGX_ERR_HIP); break; }
			unsigned long long bits = 0;
			// read on the stream that wrote it (a non-blocking caller stream does not order with the null stream)
			if (hipMemcpyAsync(&bits, c.dmax, sizeof(bits), hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
			    hipStreamSynchronize(c.stream) != hipSuccess) { fail(DGX_ERR_HIP); break; }
			std::memcpy(&mx, &bits, sizeof(mx));
			if (!(mx == mx) || mx > 1e300) { fail(DGX_ERR_NONFINITE); break; }
			if (it == 0) max_rhs0 = mx;
			if (it < DGX_HISTORY_MAX) st->history[it] = mx;
			st->newton_steps = it;
			st->max_rhs = mx;
			st->max_rhs0 = max_rhs0;
			if (mx < prm->exit_tol || (prm->exit_ratio > 0.0 && mx / max_rhs0 < prm->exit_ratio)) {
				st->converged = 1;
				break;
			}
			double cfl = prm->cfl_initial * (max_rhs0 / mx < 1.0 ? 1.0 : max_rhs0 / mx);
			if (prm->cfl_max > 0.0 && cfl > prm->cfl_max) cfl = prm->cfl_max;
			if (dgx::launch_cfl_dt(c, u, cfl, inv_dt) != hipSuccess) { fail(DGX_ERR_HIP); break; }
			// b = -r
			if (gm.blas.axpby(-1.0, r, 0.0, r, n) != hipSuccess) { fail(DGX_ERR_HIP); break; }
			if (P->setup(c, u, inv_dt) != hipSuccess) { fail(DGX_ERR_HIP); break; }
			int its = 0;
			bool conv = false;
			const int first_max = Pfb ? std::min(prm->learned_fallback_iters, prm->gmres_max_iter) : prm->gmres_max_iter;
			if (gm.solve(c, u, inv_dt, r, *P, prm->gmres_rtol, first_max, &its, &conv) != hipSuccess) {
				fail(DGX_ERR_HIP);
				break;
			}
			int step_its = its;
			if (Pfb && !conv) {
				// the learned preconditioner did not get there: exact block-Jacobi for this step
				st->fallback_steps += 1;
				if (Pfb->setup(c, u, inv_dt) != hipSuccess ||
				    gm.solve(c, u, inv_dt, r, *Pfb, prm->gmres_rtol, prm->gmres_max_iter, &its, &conv) != hipSuccess) {
					fail(DGX_ERR_HIP);
					break;
				}
				step_its += its;
			}
			its = step_its;
			if (it < DGX_HISTORY_MAX) st->gmres_per_step[it] = its;
			st->gmres_iters += its;
			if (!conv) st->gmres_not_converged += 1;  // KSP DIVERGED_ITS: warn and continue
			// s += relax * x, then the positivity limiter on every element (update_coef_s_v,
			// solve_implicit.c:723-732; enforce_positivity_highorder, solve.c:187-233)
			if (dgx::launch_fill_i32(c.stream, status, 1, 0) != hipSuccess ||
			    dgx::launch_relaxed_update(c, u, gm.x, status) != hipSuccess ||
			    dgx::launch_positivity(c, u, status) != hipSuccess) { fail(DGX_ERR_HIP); break; }
			int bad = 0;
			if (hipMemcpyAsync(&bad, status, sizeof(int), hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
			    hipStreamSynchronize(c.stream) != hipSuccess) { fail(DGX_ERR_HIP); break; }
			if (bad) { fail(DGX_ERR_POSITIVITY); break; }
			st->newton_steps = it + 1;
		}
	} while (0);
	(void)hipStreamSynchronize(c.stream);
	if (prm->precond == DGX_PC_BLOCK_JACOBI && P) st->captured = static_cast<dgx::PrecondBlockJacobi*>(P)->cap_count;
	delete P;
	delete Pfb;
	gm.release();
	if (r) (void)hipFree(r);
	if (inv_dt) (void)hipFree(inv_dt);
	if (status) (void)hipFree(status);
	st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
	if (ret == DGX_OK && !st->converged) ret = DGX_ERR_NOT_CONVERGED;
	return ret;
}
