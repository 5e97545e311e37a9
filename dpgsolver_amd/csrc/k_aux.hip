// k_aux.hip -- geometry, mass inverse, positivity, RK stages, norms, batched block LU (gfx950).
#include <hip/hip_runtime.h>

#include "dgx_device.h"
#include "dgx_internal.h"

namespace dgx {

// node coordinate helpers (r fastest)
template <int D, int N>
__device__ inline void coords(int n, int* c)
{
	c[0] = n % N;
	c[1] = (D > 1) ? (n / N) % N : 0;
	c[2] = (D > 2) ? n / (N * N) : 0;
}
template <int D, int N>
__device__ inline int with_coord(int n, int k, int j)
{
	int c[3];
	coords<D, N>(n, c);
	c[k] = j;
	return c[0] + N * (c[1] + N * c[2]);
}
// pointwise 1-D operator application along k at node n: sum_j M[c_k][j] buf[n with c_k = j]
template <int D, int N>
__device__ inline double apply1d(const double* M, const double* buf, int n, int k)
{
	int c[3];
	coords<D, N>(n, c);
	double acc = 0.0;
	for (int j = 0; j < N; ++j) acc += M[c[k] * N + j] * buf[with_coord<D, N>(n, k, j)];
	return acc;
}

// ------------------------------------------------------------------------------- geometry
// Volume metrics at the GLL metric nodes (2-D cofactors geometry_T.c:546-590; 3-D symmetric
// curl form deprecated/src/setup_geom_factors.c:157-243), interpolated to the GL volume
// cubature (geometry_T.c:734-801); det J from the Jacobian at the volume cubature.
template <int D, int N>
__global__ void __launch_bounds__(256) k_geometry(const double* __restrict__ xyz, const Ops1D<double>* ops, int E,
                                                  double* metrics_vm, double* metrics_vc, double* detJ_vc)
{
	constexpr int Np = ipow(N, D);
	__shared__ double X[D * Np], JAC[D * D * Np], MV[D * D * Np], T1[Np], T2[Np], V[3 * Np];
	__shared__ double sDg[N * N], sI[N * N];
	const int e = blockIdx.x, n = threadIdx.x;
	for (int i = threadIdx.x; i < N * N; i += blockDim.x) { sDg[i] = ops->Dg[i]; sI[i] = ops->I[i]; }
	for (int i = threadIdx.x; i < D * Np; i += blockDim.x) X[i] = xyz[(size_t)e * D * Np + i];
	__syncthreads();
	if (n < Np)
		for (int i = 0; i < D; ++i)
			for (int k = 0; k < D; ++k) JAC[(i * D + k) * Np + n] = apply1d<D, N>(sDg, X + i * Np, n, k);
	__syncthreads();
	if constexpr (D == 2) {
		if (n < Np) {
			const double x_r = JAC[0 * Np + n], x_s = JAC[1 * Np + n], y_r = JAC[2 * Np + n], y_s = JAC[3 * Np + n];
			MV[(0 + 2 * 0) * Np + n] = y_s;
			MV[(0 + 2 * 1) * Np + n] = -y_r;
			MV[(1 + 2 * 0) * Np + n] = -x_s;
			MV[(1 + 2 * 1) * Np + n] = x_r;
		}
		__syncthreads();
	} else {
		const int pairs[3][2] = { { 1, 2 }, { 2, 0 }, { 0, 1 } };
		const int cyc[3][2] = { { 1, 2 }, { 2, 0 }, { 0, 1 } };
		double cu[2][3];
		for (int row = 0; row < 3; ++row) {
			for (int tt = 0; tt < 2; ++tt) {
				const int a = tt == 0 ? pairs[row][0] : pairs[row][1];
				const int b = tt == 0 ? pairs[row][1] : pairs[row][0];
				if (n < Np)
					for (int k = 0; k < 3; ++k) V[k * Np + n] = X[a * Np + n] * JAC[(b * 3 + k) * Np + n];
				__syncthreads();
				if (n < Np)
					for (int col = 0; col < 3; ++col) {
						const int k1 = cyc[col][0], k2 = cyc[col][1];
						cu[tt][col] = apply1d<D, N>(sDg, V + k2 * Np, n, k1) - apply1d<D, N>(sDg, V + k1 * Np, n, k2);
					}
				__syncthreads();
			}
			if (n < Np)
				for (int col = 0; col < 3; ++col) MV[(row + 3 * col) * Np + n] = 0.5 * (cu[0][col] - cu[1][col]);
		}
		__syncthreads();
	}
	if (n < Np)
		for (int q = 0; q < D * D; ++q) metrics_vm[((size_t)e * D * D + q) * Np + n] = MV[q * Np + n];
	// interpolate metrics and the Jacobian to the GL nodes (D pointwise passes, T1/T2 ping-pong)
	for (int comp = 0; comp < 2 * D * D; ++comp) {
		const double* src = comp < D * D ? MV + comp * Np : JAC + (comp - D * D) * Np;
		double* cur = (double*)src;
		for (int k = 0; k < D; ++k) {
			double* dst = (k % 2 == 0) ? T1 : T2;
			double val = 0.0;
			if (n < Np) val = apply1d<D, N>(sI, cur, n, k);
			__syncthreads();
			if (n < Np) dst[n] = val;
			__syncthreads();
			cur = dst;
		}
		if (comp < D * D) {
			if (n < Np) metrics_vc[((size_t)e * D * D + comp) * Np + n] = cur[n];
		} else if (n < Np) {
			MV[(comp - D * D) * Np + n] = cur[n];  // MV reused: Jacobian at vc
		}
		__syncthreads();
	}
	if (n < Np) {
#define JV(i, k) MV[((i) * D + (k)) * Np + n]
		double dj;
		if constexpr (D == 2) dj = JV(0, 0) * JV(1, 1) - JV(0, 1) * JV(1, 0);
		else
			dj = JV(0, 0) * (JV(1, 1) * JV(2, 2) - JV(1, 2) * JV(2, 1))
			   - JV(0, 1) * (JV(1, 0) * JV(2, 2) - JV(1, 2) * JV(2, 0))
			   + JV(0, 2) * (JV(1, 0) * JV(2, 1) - JV(1, 1) * JV(2, 0));
#undef JV
		detJ_vc[(size_t)e * Np + n] = dj;
	}
}

// face n*J_f at the face cubature from the LEFT element's metrics (geometry_T.c:274-325, 412-445)
template <int D, int N>
__global__ void k_face_geometry(const int* __restrict__ faces, const double* __restrict__ metrics_vm,
                                const Ops1D<double>* ops, int F, double* nJ_fc)
{
	constexpr int Np = ipow(N, D), NF = ipow(N, D - 1);
	const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (gid >= (long)F * NF) return;
	const int f = (int)(gid / NF), i = (int)(gid % NF);
	const int eL = faces[6 * f], lf = faces[6 * f + 1];
	const int fd = lf >> 1, fix = (lf & 1) ? N - 1 : 0;
	const double sgn = (lf & 1) ? 1.0 : -1.0;
	const int qa = i % N, qb = (D == 3) ? i / N : 0;
	const double* I = ops->I;
	for (int comp = 0; comp < D; ++comp) {
		// m[comp + D*fd] on face lf interpolated GLL -> GL along the face coordinates
		const double* m = metrics_vm + ((size_t)eL * D * D + comp + D * fd) * Np;
		double acc = 0.0;
		if constexpr (D == 2) {
			for (int a = 0; a < N; ++a) {
				const int node = (fd == 0) ? fix + N * a : a + N * fix;
				acc += I[qa * N + a] * m[node];
			}
		} else {
			for (int b = 0; b < N; ++b) {
				double inner = 0.0;
				for (int a = 0; a < N; ++a) {
					int node;
					if (fd == 0) node = fix + N * (a + N * b);
					else if (fd == 1) node = a + N * (fix + N * b);
					else node = a + N * (b + N * fix);
					inner += I[qa * N + a] * m[node];
				}
				acc += I[qb * N + b] * inner;
			}
		}
		nJ_fc[((size_t)f * NF + i) * D + comp] = sgn * acc;
	}
}

hipError_t launch_geometry(Ctx& c)
{
	double* mvm = nullptr;
	DGX_HIP_CHECK(hipMalloc(&mvm, sizeof(double) * (size_t)c.E * c.d * c.d * c.Np));
	hipError_t err = hipErrorInvalidValue;
	const int NF = c.Nfc;
	const int fgrid = (int)(((long)c.F * NF + 255) / 256);
#define DGX_GEOM(DD, NN)                                                                                    \
	if (c.d == DD && c.N == NN) {                                                                           \
		hipLaunchKernelGGL((k_geometry<DD, NN>), dim3(c.E), dim3(ipow(NN, DD) < 64 ? 64 : ipow(NN, DD)), 0,  \
		                   c.stream, c.xyz, c.ops64, c.E, mvm, c.metrics_vc, c.detJ_vc);                    \
		err = hipGetLastError();                                                                             \
		if (err == hipSuccess && c.F > 0)                                                                    \
			hipLaunchKernelGGL((k_face_geometry<DD, NN>), dim3(fgrid), dim3(256), 0, c.stream, c.faces, mvm, \
			                   c.ops64, c.F, c.nJ_fc);                                                       \
		if (err == hipSuccess) err = hipGetLastError();                                                      \
	}
	DGX_GEOM(2, 2) DGX_GEOM(2, 3) DGX_GEOM(2, 4) DGX_GEOM(2, 5) DGX_GEOM(2, 6) DGX_GEOM(2, 7) DGX_GEOM(2, 8)
	DGX_GEOM(3, 2) DGX_GEOM(3, 3) DGX_GEOM(3, 4) DGX_GEOM(3, 5)
#undef DGX_GEOM
	hipError_t e2 = hipStreamSynchronize(c.stream);
	(void)hipFree(mvm);
	return err != hipSuccess ? err : e2;
}

// ------------------------------------------------------------------------------- mass inverse
// M = cv0^T diag(w J) cv0 with square invertible cv0 = Ibar, so M^-1 = Ibar^-1 diag(1/(wJ)) Ibar^-T
// (volume_solver_T.c:115-145 forms and inverts M densely).
template <int D, int N, int NV>
__global__ void __launch_bounds__(256) k_minv(double* r, const double* __restrict__ detJ, const Ops1D<double>* ops,
                                              int E)
{
	constexpr int Np = ipow(N, D);
	__shared__ double B0[NV * Np], B1[NV * Np], sIinv[N * N], sIinvT[N * N], sW[N];
	const int e = blockIdx.x, n = threadIdx.x;
	for (int i = threadIdx.x; i < N * N; i += blockDim.x) {
		sIinv[i] = ops->Iinv[i];
		sIinvT[i] = ops->Iinv[(i % N) * N + i / N];
	}
	for (int i = threadIdx.x; i < N; i += blockDim.x) sW[i] = ops->w[i];
	for (int i = threadIdx.x; i < NV * Np; i += blockDim.x) B0[i] = r[(size_t)e * NV * Np + i];
	__syncthreads();
	double* cur = B0;
	double* nxt = B1;
	for (int pass = 0; pass < 2 * D; ++pass) {
		const double* M = pass < D ? sIinvT : sIinv;
		const int k = pass % D;
		if (n < Np)
			for (int v = 0; v < NV; ++v) nxt[v * Np + n] = apply1d<D, N>(M, cur + v * Np, n, k);
		__syncthreads();
		if (pass == D - 1 && n < Np) {
			int c[3];
			coords<D, N>(n, c);
			double wq = sW[c[0]] * sW[c[1]];
			if (D == 3) wq *= sW[c[2]];
			const double s = 1.0 / (wq * detJ[(size_t)e * Np + n]);
			for (int v = 0; v < NV; ++v) nxt[v * Np + n] *= s;
		}
		__syncthreads();
		double* t = cur; cur = nxt; nxt = t;
	}
	for (int i = threadIdx.x; i < NV * Np; i += blockDim.x) r[(size_t)e * NV * Np + i] = cur[i];
}

// ------------------------------------------------------------------------------- positivity
// enforce_positivity_highorder (solve.c:187-233) with correct_coef (solve.c:356-374).
template <int D, int N>
__global__ void __launch_bounds__(256) k_positivity(double* u, const Ops1D<double>* ops, int E, int* status)
{
	constexpr int Np = ipow(N, D), NV = D + 2;
	__shared__ double B0[NV * Np], B1[NV * Np], sB[N * N], sBinv[N * N];
	__shared__ double sAvg[NV], sPho;
	__shared__ int sBad;
	const int e = blockIdx.x, n = threadIdx.x;
	for (int i = threadIdx.x; i < N * N; i += blockDim.x) { sB[i] = ops->B[i]; sBinv[i] = ops->Binv[i]; }
	for (int i = threadIdx.x; i < NV * Np; i += blockDim.x) B0[i] = u[(size_t)e * NV * Np + i];
	__syncthreads();
	double* cur = B0;
	double* nxt = B1;
	for (int k = 0; k < D; ++k) {
		if (n < Np)
			for (int v = 0; v < NV; ++v) nxt[v * Np + n] = apply1d<D, N>(sBinv, cur + v * Np, n, k);
		__syncthreads();
		double* t = cur; cur = nxt; nxt = t;
	}
	if (n < Np) {  // conserved -> primitive on the Bezier coefficients
		const double rho = cur[n];
		double V2 = 0.0;
		for (int k = 0; k < D; ++k) {
			cur[(1 + k) * Np + n] /= rho;
			V2 += cur[(1 + k) * Np + n] * cur[(1 + k) * Np + n];
		}
		cur[(D + 1) * Np + n] = kGM1 * (cur[(D + 1) * Np + n] - 0.5 * rho * V2);
	}
	if (threadIdx.x == 0) sBad = 0;
	__syncthreads();
	for (int vr = 0; vr < NV; vr += NV - 1) {
		if (threadIdx.x == 0) {
			const double* vd = cur + vr * Np;
			double avg = 0.0, mn = vd[0];
			for (int i = 0; i < Np; ++i) { avg += vd[i]; mn = fmin(mn, vd[i]); }
			avg /= Np;
			sPho = -1.0;
			if (avg < kEpsPhys) sBad = 1;
			else if (mn < kEpsPhys) {
				double p = (avg - kEpsPhys) / (avg - mn);
				sPho = p < 0.0 ? 0.0 : (p > 1.0 ? 1.0 : p);
				for (int v2 = 0; v2 < NV; ++v2) {
					double a2 = 0.0;
					for (int i = 0; i < Np; ++i) a2 += cur[v2 * Np + i];
					a2 /= Np;
					if ((v2 == 0 || v2 == NV - 1) && a2 < kEpsPhys) a2 = kEpsPhys;
					sAvg[v2] = a2;
				}
			}
		}
		__syncthreads();
		if (sPho >= 0.0 && n < Np)
			for (int v2 = 0; v2 < NV; ++v2) cur[v2 * Np + n] = sPho * cur[v2 * Np + n] + (1 - sPho) * sAvg[v2];
		__syncthreads();
	}
	if (n < Np) {  // primitive -> conserved
		const double rho = cur[n];
		double V2 = 0.0;
		for (int k = 0; k < D; ++k) V2 += cur[(1 + k) * Np + n] * cur[(1 + k) * Np + n];
		for (int k = 0; k < D; ++k) cur[(1 + k) * Np + n] *= rho;
		cur[(D + 1) * Np + n] = cur[(D + 1) * Np + n] / kGM1 + 0.5 * rho * V2;
	}
	__syncthreads();
	for (int k = 0; k < D; ++k) {
		if (n < Np)
			for (int v = 0; v < NV; ++v) nxt[v * Np + n] = apply1d<D, N>(sB, cur + v * Np, n, k);
		__syncthreads();
		double* t = cur; cur = nxt; nxt = t;
	}
	for (int i = threadIdx.x; i < NV * Np; i += blockDim.x) u[(size_t)e * NV * Np + i] = cur[i];
	if (threadIdx.x == 0 && sBad) atomicOr(status, 1);
}

#define DGX_PER_ELEM(KNAME, ...)                                                                       \
	switch (c.d * 16 + c.N) {                                                                          \
	case 2 * 16 + 2: hipLaunchKernelGGL((KNAME<2, 2 __VA_ARGS__>), dim3(c.E), dim3(64), 0, c.stream, ARGS); break;  \
	case 2 * 16 + 3: hipLaunchKernelGGL((KNAME<2, 3 __VA_ARGS__>), dim3(c.E), dim3(64), 0, c.stream, ARGS); break;  \
	case 2 * 16 + 4: hipLaunchKernelGGL((KNAME<2, 4 __VA_ARGS__>), dim3(c.E), dim3(64), 0, c.stream, ARGS); break;  \
	case 2 * 16 + 5: hipLaunchKernelGGL((KNAME<2, 5 __VA_ARGS__>), dim3(c.E), dim3(64), 0, c.stream, ARGS); break;  \
	case 2 * 16 + 6: hipLaunchKernelGGL((KNAME<2, 6 __VA_ARGS__>), dim3(c.E), dim3(64), 0, c.stream, ARGS); break;  \
	case 3 * 16 + 2: hipLaunchKernelGGL((KNAME<3, 2 __VA_ARGS__>), dim3(c.E), dim3(64), 0, c.stream, ARGS); break;  \
	case 3 * 16 + 3: hipLaunchKernelGGL((KNAME<3, 3 __VA_ARGS__>), dim3(c.E), dim3(64), 0, c.stream, ARGS); break;  \
	case 3 * 16 + 4: hipLaunchKernelGGL((KNAME<3, 4 __VA_ARGS__>), dim3(c.E), dim3(64), 0, c.stream, ARGS); break;  \
	case 3 * 16 + 5: hipLaunchKernelGGL((KNAME<3, 5 __VA_ARGS__>), dim3(c.E), dim3(128), 0, c.stream, ARGS); break; \
	default: return hipErrorInvalidValue;                                                              \
	}

hipError_t launch_minv(Ctx& c, double* r)
{
	if (c.E == 0) return hipSuccess;
#define ARGS r, c.detJ_vc, c.ops64, c.E
	if (c.nv == 1) { DGX_PER_ELEM(k_minv, , 1) }
	else if (c.d == 2) { DGX_PER_ELEM(k_minv, , 4) }
	else { DGX_PER_ELEM(k_minv, , 5) }
#undef ARGS
	return hipGetLastError();
}

hipError_t launch_positivity(Ctx& c, double* u, int* status)
{
	if (c.E == 0) return hipSuccess;
#define ARGS u, c.ops64, c.E, status
	DGX_PER_ELEM(k_positivity, )
#undef ARGS
	return hipGetLastError();
}

// ------------------------------------------------------------------------------- elementwise
__global__ void k_maxabs(const double* __restrict__ x, size_t n, unsigned long long* out)
{
	double m = 0.0;
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
		const double a = fabs(x[i]);
		m = (a > m || a != a) ? a : m;  // NaN propagates as the largest bit pattern
	}
	for (int off = 32; off > 0; off >>= 1) {
		const double o = __shfl_down(m, off);
		m = (o > m || o != o) ? o : m;
	}
	if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

hipError_t launch_maxabs(Ctx& c, const double* r, size_t n, unsigned long long* dmax)
{
	const int grid = (int)((n + 255) / 256 < 2048 ? (n + 255) / 256 : 2048);
	if (grid == 0) return hipSuccess;
	hipLaunchKernelGGL(k_maxabs, dim3(grid), dim3(256), 0, c.stream, r, n, dmax);
	return hipGetLastError();
}

__global__ void k_f64_to_f32(const double* __restrict__ s, float* __restrict__ d, size_t n)
{
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		d[i] = (float)s[i];
}
__global__ void k_f32_to_f64(const float* __restrict__ s, double* __restrict__ d, size_t n)
{
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		d[i] = (double)s[i];
}
hipError_t launch_convert(const double* src, float* dst, size_t n, hipStream_t s)
{
	const int grid = (int)((n + 255) / 256 < 8192 ? (n + 255) / 256 : 8192);
	if (grid) hipLaunchKernelGGL(k_f64_to_f32, dim3(grid), dim3(256), 0, s, src, dst, n);
	return hipGetLastError();
}
hipError_t launch_convert_back(const float* src, double* dst, size_t n, hipStream_t s)
{
	const int grid = (int)((n + 255) / 256 < 8192 ? (n + 255) / 256 : 8192);
	if (grid) hipLaunchKernelGGL(k_f32_to_f64, dim3(grid), dim3(256), 0, s, src, dst, n);
	return hipGetLastError();
}

// SSP-RK3 stage arithmetic (solve_explicit.c:276-295)
__global__ void k_rk_stage(double* __restrict__ u, double* __restrict__ u0, const double* __restrict__ r, double dt,
                           int stage, size_t n)
{
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
		if (stage == 0) {
			u0[i] = u[i];
			u[i] += dt * r[i];
		} else if (stage == 1) {
			u[i] = (1.0 / 4.0) * (3.0 * u0[i] + u[i] + dt * r[i]);
		} else {
			u[i] = (1.0 / 3.0) * (u0[i] + 2.0 * u[i] + 2.0 * dt * r[i]);
		}
	}
}
hipError_t launch_rk_stage_buf(Ctx& c, double* u, double* u0, const double* r, double dt, int stage)
{
	const size_t n = (size_t)c.E * c.Np * c.nv;
	const int grid = (int)((n + 255) / 256 < 8192 ? (n + 255) / 256 : 8192);
	if (grid) hipLaunchKernelGGL(k_rk_stage, dim3(grid), dim3(256), 0, c.stream, u, u0, r, dt, stage, n);
	return hipGetLastError();
}
hipError_t launch_rk_stage(Ctx& c, double* u, const double* r, double dt, int stage)
{
	return launch_rk_stage_buf(c, u, c.u0, r, dt, stage);
}
hipError_t launch_rk_stage_n(hipStream_t st, double* u, double* u0, const double* r, double dt, int stage, size_t n)
{
	const int grid = (int)((n + 255) / 256 < 8192 ? (n + 255) / 256 : 8192);
	if (grid) hipLaunchKernelGGL(k_rk_stage, dim3(grid), dim3(256), 0, st, u, u0, r, dt, stage, n);
	return hipGetLastError();
}

// ------------------------------------------------------------------------------- mass blocks
// blocks[e] -= scale * (M_e (x) I_nv), M_e[i][j] = sum_q Ibar[q][i] w_q J_q Ibar[q][j]
// (the CFL-ramping term A += -M/dt, solve_dg.c:223-248, 284-319)
template <int D, int N>
__global__ void k_sub_mass(double* blocks, const double* __restrict__ detJ, const Ops1D<double>* ops, int E, int nv,
                           const double* scale_e, double scale_c)
{
	const double scale = scale_e ? scale_e[blockIdx.x] : scale_c;
	constexpr int Np = ipow(N, D);
	__shared__ double sI[N * N], sW[N], wJ[Np];
	const int e = blockIdx.x;
	for (int i = threadIdx.x; i < N * N; i += blockDim.x) sI[i] = ops->I[i];
	for (int i = threadIdx.x; i < N; i += blockDim.x) sW[i] = ops->w[i];
	__syncthreads();
	for (int q = threadIdx.x; q < Np; q += blockDim.x) {
		int c[3];
		coords<D, N>(q, c);
		double w = sW[c[0]] * sW[c[1]];
		if (D == 3) w *= sW[c[2]];
		wJ[q] = w * detJ[(size_t)e * Np + q];
	}
	__syncthreads();
	const int nb = Np * nv;
	for (int ij = threadIdx.x; ij < Np * Np; ij += blockDim.x) {
		const int i = ij / Np, j = ij % Np;
		int ci[3], cj[3];
		coords<D, N>(i, ci);
		coords<D, N>(j, cj);
		double acc = 0.0;
		for (int q = 0; q < Np; ++q) {
			int cq[3];
			coords<D, N>(q, cq);
			double bi = 1.0, bj = 1.0;
			for (int k = 0; k < D; ++k) { bi *= sI[cq[k] * N + ci[k]]; bj *= sI[cq[k] * N + cj[k]]; }
			acc += bi * wJ[q] * bj;
		}
		for (int v = 0; v < nv; ++v)
			blocks[(size_t)e * nb * nb + (size_t)(v * Np + i) * nb + (v * Np + j)] -= scale * acc;
	}
}
hipError_t launch_sub_mass(Ctx& c, double* blocks, const double* scale_e, double scale)
{
	if (c.E == 0) return hipSuccess;
#define ARGS blocks, c.detJ_vc, c.ops64, c.E, c.nv, scale_e, scale
	DGX_PER_ELEM(k_sub_mass, )
#undef ARGS
	return hipGetLastError();
}

// ------------------------------------------------------------------------------- batched block LU
// In-place LU with partial pivoting of one nb x nb block per workgroup (nb <= 128), LDS resident.
__global__ void __launch_bounds__(256) k_block_lu(double* blocks, int* piv, int nb)
{
	extern __shared__ double A[];  // nb*nb
	__shared__ int sp;
	const int e = blockIdx.x;
	double* g = blocks + (size_t)e * nb * nb;
	for (int i = threadIdx.x; i < nb * nb; i += blockDim.x) A[i] = g[i];
	__syncthreads();
	for (int k = 0; k < nb; ++k) {
		if (threadIdx.x == 0) {
			int p = k;
			double best = fabs(A[k * nb + k]);
			for (int i = k + 1; i < nb; ++i)
				if (fabs(A[i * nb + k]) > best) { best = fabs(A[i * nb + k]); p = i; }
			sp = p;
			piv[(size_t)e * nb + k] = p;
		}
		__syncthreads();
		const int p = sp;
		if (p != k)
			for (int j = threadIdx.x; j < nb; j += blockDim.x) {
				const double t = A[k * nb + j];
				A[k * nb + j] = A[p * nb + j];
				A[p * nb + j] = t;
			}
		__syncthreads();
		const double inv = 1.0 / A[k * nb + k];
		for (int i = k + 1 + threadIdx.x; i < nb; i += blockDim.x) A[i * nb + k] *= inv;
		__syncthreads();
		const int m = nb - k - 1;
		for (int ij = threadIdx.x; ij < m * m; ij += blockDim.x) {
			const int i = k + 1 + ij / m, j = k + 1 + ij % m;
			A[i * nb + j] -= A[i * nb + k] * A[k * nb + j];
		}
		__syncthreads();
	}
	for (int i = threadIdx.x; i < nb * nb; i += blockDim.x) g[i] = A[i];
}

// x_e <- U^-1 L^-1 P x_e, one wavefront per block
__global__ void __launch_bounds__(64) k_block_lu_solve(const double* __restrict__ lu, const int* __restrict__ piv,
                                                       double* x, int nb)
{
	extern __shared__ double b[];
	const int e = blockIdx.x, l = threadIdx.x;
	const double* A = lu + (size_t)e * nb * nb;
	double* xe = x + (size_t)e * nb;
	for (int i = l; i < nb; i += 64) b[i] = xe[i];
	__syncthreads();
	if (l == 0)
		for (int k = 0; k < nb; ++k) {
			const int p = piv[(size_t)e * nb + k];
			if (p != k) { const double t = b[k]; b[k] = b[p]; b[p] = t; }
		}
	__syncthreads();
	for (int k = 0; k < nb; ++k) {  // forward, unit lower
		const double bk = b[k];
		for (int i = k + 1 + l; i < nb; i += 64) b[i] -= A[i * nb + k] * bk;
		__syncthreads();
	}
	for (int k = nb - 1; k >= 0; --k) {  // backward
		if (l == 0) b[k] /= A[k * nb + k];
		__syncthreads();
		const double bk = b[k];
		for (int i = l; i < k; i += 64) b[i] -= A[i * nb + k] * bk;
		__syncthreads();
	}
	for (int i = l; i < nb; i += 64) xe[i] = b[i];
}

// In-place inverse of every block by Gauss-Jordan elimination with partial pivoting, stored transposed
// (out[j * nb + i] = (A^-1)[i][j]) so that the apply kernel reads it coalesced.
__global__ void __launch_bounds__(256) k_block_inverse(double* blocks, int nb, int* status)
{
	extern __shared__ double A[];  // nb*nb, then the pivot column (nb)
	double* col = A + nb * nb;
	__shared__ int sp, sperm[kMaxInvBlock];
	const int e = blockIdx.x;
	double* g = blocks + (size_t)e * nb * nb;
	for (int i = threadIdx.x; i < nb * nb; i += blockDim.x) A[i] = g[i];
	__syncthreads();
	for (int k = 0; k < nb; ++k) {
		if (threadIdx.x == 0) {
			int p = k;
			double best = fabs(A[k * nb + k]);
			for (int i = k + 1; i < nb; ++i)
				if (fabs(A[i * nb + k]) > best) { best = fabs(A[i * nb + k]); p = i; }
			sp = p;
			sperm[k] = p;
			if (!(best > 0.0) && status) atomicOr(status, 1);
		}
		__syncthreads();
		const int p = sp;
		if (p != k)
			for (int j = threadIdx.x; j < nb; j += blockDim.x) {
				const double t = A[k * nb + j];
				A[k * nb + j] = A[p * nb + j];
				A[p * nb + j] = t;
			}
		__syncthreads();
		const double inv = 1.0 / A[k * nb + k];
		for (int i = threadIdx.x; i < nb; i += blockDim.x) col[i] = A[i * nb + k];
		__syncthreads();
		for (int j = threadIdx.x; j < nb; j += blockDim.x) A[k * nb + j] = (j == k ? 1.0 : A[k * nb + j]) * inv;
		__syncthreads();
		for (int ij = threadIdx.x; ij < nb * nb; ij += blockDim.x) {
			const int i = ij / nb, j = ij % nb;
			if (i == k) continue;
			if (j == k) A[ij] = -col[i] * A[k * nb + k];
			else A[ij] -= col[i] * A[k * nb + j];
		}
		__syncthreads();
	}
	// undo the row interchanges as column interchanges, last first
	for (int k = nb - 1; k >= 0; --k) {
		const int p = sperm[k];
		if (p != k)
			for (int i = threadIdx.x; i < nb; i += blockDim.x) {
				const double t = A[i * nb + k];
				A[i * nb + k] = A[i * nb + p];
				A[i * nb + p] = t;
			}
		__syncthreads();
	}
	for (int ij = threadIdx.x; ij < nb * nb; ij += blockDim.x) {
		const int i = ij / nb, j = ij % nb;
		g[(size_t)j * nb + i] = A[ij];
	}
}

// z_e = A_e^-1 v_e with the transposed inverse from k_block_inverse (one wavefront per element)
__global__ void __launch_bounds__(64) k_block_inv_apply(const double* __restrict__ invT, const double* __restrict__ v,
                                                        double* z, int nb)
{
	extern __shared__ double sv[];
	const int e = blockIdx.x, l = threadIdx.x;
	const double* M = invT + (size_t)e * nb * nb;
	for (int i = l; i < nb; i += 64) sv[i] = v[(size_t)e * nb + i];
	__syncthreads();
	for (int i = l; i < nb; i += 64) {
		double acc = 0.0;
		for (int k = 0; k < nb; ++k) acc += M[(size_t)k * nb + i] * sv[k];
		z[(size_t)e * nb + i] = acc;
	}
}

hipError_t launch_block_inverse(hipStream_t s, double* blocks, int E, int nb, int* status)
{
	if (E == 0) return hipSuccess;
	if (nb <= 0 || nb > kMaxInvBlock) return hipErrorInvalidValue;
	const size_t lds = sizeof(double) * (nb * nb + nb);
	if (lds > 65536) {   // beyond 64 KB the launch asks for the larger dynamic LDS once per process
		static hipError_t attr = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_block_inverse),
		                                             hipFuncAttributeMaxDynamicSharedMemorySize,
		                                             (int)(sizeof(double) * (kMaxInvBlock * kMaxInvBlock + kMaxInvBlock)));
		if (attr != hipSuccess) return attr;
	}
	hipLaunchKernelGGL(k_block_inverse, dim3(E), dim3(256), lds, s, blocks, nb, status);
	return hipGetLastError();
}
hipError_t launch_block_inv_apply(hipStream_t s, const double* invT, const double* v, double* z, int E, int nb)
{
	if (E == 0) return hipSuccess;
	hipLaunchKernelGGL(k_block_inv_apply, dim3(E), dim3(64), sizeof(double) * nb, s, invT, v, z, nb);
	return hipGetLastError();
}

// Face-node map: volume node of local face lf at face position i (faces r-, r+, s-, s+[, t-, t+]; on a HEX face
// i = a + N b, a and b the two remaining directions in increasing order)
__device__ inline int face_node_dd(int d, int lf, int i, int N)
{
	const int fd = lf >> 1, fix = (lf & 1) ? N - 1 : 0;
	if (d == 2) return fd == 0 ? fix + N * i : i + N * fix;
	const int a = i % N, b = i / N;
	if (fd == 0) return fix + N * (a + N * b);
	if (fd == 1) return a + N * (fix + N * b);
	return a + N * (b + N * fix);
}

// Compact transposed face blocks (nfv = nv N^(d-1) face-node rows and columns, [col][row], as the MODE_OFFD
// launches store them) scattered into full row-major nb x nb blocks, zero off the face nodes and on boundary
// faces.  One workgroup per (element, local face).
__global__ void k_expand_face_blocks(const double* __restrict__ cT, const ElemFace* __restrict__ ef, double* full,
                                     int d, int N, int nv)
{
	const int NF = (d == 2) ? N : N * N, Np = NF * N, nb = Np * nv, nfv = nv * NF;
	const size_t b = blockIdx.x;
	const int lf = (int)(b % (2 * d));
	const ElemFace f = ef[b];
	double* out = full + b * nb * nb;
	for (int k = threadIdx.x; k < nb * nb; k += blockDim.x) out[k] = 0.0;
	__syncthreads();
	if (f.en < 0) return;
	const int fn = f.packed & 15;
	for (int k = threadIdx.x; k < nfv * nfv; k += blockDim.x) {
		const int col = k / nfv, row = k - col * nfv;
		const int v = row / NF, w = col / NF;
		out[(size_t)(v * Np + face_node_dd(d, lf, row - v * NF, N)) * nb + w * Np + face_node_dd(d, fn, col - w * NF, N)] =
			cT[b * nfv * nfv + k];
	}
}

hipError_t launch_expand_face_blocks(Ctx& c, const double* cT, double* full)
{
	if (c.d != 2 && c.d != 3) return hipErrorInvalidValue;
	if (c.E == 0) return hipSuccess;
	hipLaunchKernelGGL(k_expand_face_blocks, dim3((unsigned)(c.E * 2 * c.d)), dim3(256), 0, c.stream, cT, c.elem_faces,
	                   full, c.d, c.N, c.nv);
	return hipGetLastError();
}

// values[slot[e]] = (or +=) src[e * stride + off], nb2 values per block, for every element with slot[e] >= 0
__global__ void __launch_bounds__(256) k_bsr_place(const double* __restrict__ src, size_t stride, size_t off,
                                                   const int* __restrict__ slot, double* vals, int nb2, int add)
{
	const int e = blockIdx.x, sl = slot[e];
	if (sl < 0) return;
	const double* a = src + (size_t)e * stride + off;
	double* b = vals + (size_t)sl * nb2;
	for (int i = threadIdx.x; i < nb2; i += blockDim.x) b[i] = add ? b[i] + a[i] : a[i];
}
hipError_t launch_bsr_place(hipStream_t s, const double* src, size_t stride, size_t off, const int* slot, double* vals,
                            int E, int nb2, bool add)
{
	if (E == 0) return hipSuccess;
	hipLaunchKernelGGL(k_bsr_place, dim3(E), dim3(256), 0, s, src, stride, off, slot, vals, nb2, add ? 1 : 0);
	return hipGetLastError();
}

hipError_t launch_block_lu(hipStream_t s, double* blocks, int* piv, int E, int nb)
{
	if (nb <= 0 || nb > kMaxLuBlock) return hipErrorInvalidValue;
	if (E == 0) return hipSuccess;
	const size_t lds = sizeof(double) * nb * nb;
	if (lds > 65536) {   // beyond 64 KB the launch asks for the larger dynamic LDS once per process
		static hipError_t attr = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_block_lu),
		                                             hipFuncAttributeMaxDynamicSharedMemorySize,
		                                             (int)(sizeof(double) * kMaxLuBlock * kMaxLuBlock));
		if (attr != hipSuccess) return attr;
	}
	hipLaunchKernelGGL(k_block_lu, dim3(E), dim3(256), lds, s, blocks, piv, nb);
	return hipGetLastError();
}
hipError_t launch_block_lu_solve(hipStream_t s, const double* lu, const int* piv, double* x, int E, int nb)
{
	if (E == 0) return hipSuccess;
	hipLaunchKernelGGL(k_block_lu_solve, dim3(E), dim3(64), sizeof(double) * nb, s, lu, piv, x, nb);
	return hipGetLastError();
}

}  // namespace dgx

namespace dgx {
// small fill kernels (no hipMemsetAsync: callers may capture these launches in a graph)
__global__ void k_fill_u64(unsigned long long* p, int n, unsigned long long v)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p[i] = v;
}
__global__ void k_fill_i32(int* p, int n, int v)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p[i] = v;
}
hipError_t launch_fill_u64(hipStream_t s, unsigned long long* p, int n, unsigned long long v)
{
	hipLaunchKernelGGL(k_fill_u64, dim3((n + 63) / 64), dim3(64), 0, s, p, n, v);
	return hipGetLastError();
}
hipError_t launch_fill_i32(hipStream_t s, int* p, int n, int v)
{
	hipLaunchKernelGGL(k_fill_i32, dim3((n + 63) / 64), dim3(64), 0, s, p, n, v);
	return hipGetLastError();
}
}  // namespace dgx
