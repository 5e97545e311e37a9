// dgx_device.h -- device-side scalar types and physics shared by the gfx950 kernels.
//
// Physics restated from the reference (paths relative to the DPGSolver source tree):
//   Euler flux                 src/simulation/test_case/flux/flux_euler_T.c:71-143, 259-297
//   Lax-Friedrichs             src/simulation/solvers/numerical_flux/numerical_flux_euler_T.c:60-327
//   Roe-Pike + entropy fix     numerical_flux_euler_T.c:470-778 (Qu et al. 2015 eq. 35)
//   upwind                     numerical_flux_advection_T.c:39-124
//   Riemann / slip-wall states src/simulation/test_case/boundary/boundary_euler_T.c:99-180, 494-530
// The arithmetic order follows the reference's expressions so that the FP64 path agrees
// with the oracle to rounding.
#pragma once

#include <hip/hip_runtime.h>

#include "dgx_internal.h"

namespace dgx {

// Forward-mode dual number: exact directional derivatives of the residual
// (the analytic Jacobians of the reference, flux_euler_T.c:298-...,
// numerical_flux_euler_T.c:329-468, 780-1648, follow from the same chain rule).
struct Dual {
	double v, d;
	Dual() = default;
	__host__ __device__ Dual(double a) : v(a), d(0.0) {}
	__host__ __device__ Dual(double a, double b) : v(a), d(b) {}
};
__device__ inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__device__ inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__device__ inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__device__ inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ inline Dual operator/(Dual a, Dual b)
{
	const double q = a.v / b.v;
	return Dual(q, (a.d - q * b.d) / b.v);
}
__device__ inline Dual& operator+=(Dual& a, Dual b) { a = a + b; return a; }
__device__ inline Dual& operator-=(Dual& a, Dual b) { a = a - b; return a; }
__device__ inline Dual& operator*=(Dual& a, Dual b) { a = a * b; return a; }

// The same in FP32 (value and derivative): the inner Jacobian-vector product of the implicit solve, read from and
// written to FP64 vectors.
struct Dual32 {
	float v, d;
	Dual32() = default;
	__host__ __device__ Dual32(double a) : v((float)a), d(0.0f) {}
	__host__ __device__ Dual32(double a, double b) : v((float)a), d((float)b) {}
	__host__ __device__ Dual32(float a, float b) : v(a), d(b) {}
};
__device__ inline Dual32 operator+(Dual32 a, Dual32 b) { return Dual32(a.v + b.v, a.d + b.d); }
__device__ inline Dual32 operator-(Dual32 a, Dual32 b) { return Dual32(a.v - b.v, a.d - b.d); }
__device__ inline Dual32 operator-(Dual32 a) { return Dual32(-a.v, -a.d); }
__device__ inline Dual32 operator*(Dual32 a, Dual32 b) { return Dual32(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ inline Dual32 operator/(Dual32 a, Dual32 b)
{
	const float q = a.v / b.v;
	return Dual32(q, (a.d - q * b.d) / b.v);
}
__device__ inline Dual32& operator+=(Dual32& a, Dual32 b) { a = a + b; return a; }
__device__ inline Dual32& operator-=(Dual32& a, Dual32 b) { a = a - b; return a; }
__device__ inline Dual32& operator*=(Dual32& a, Dual32 b) { a = a * b; return a; }

__device__ inline double rv(double a) { return a; }
__device__ inline float rv(float a) { return a; }
__device__ inline double rv(Dual a) { return a.v; }
__device__ inline float rv(Dual32 a) { return a.v; }

__device__ inline double tsqrt(double a) { return sqrt(a); }
// FP32: the hardware square root and reciprocal (1 ulp), not the IEEE-rounded sequences
__device__ inline float tsqrt(float a) { return __builtin_amdgcn_sqrtf(a); }
__device__ inline Dual tsqrt(Dual a)
{
	const double s = sqrt(a.v);
	return Dual(s, a.d / (2.0 * s));
}
__device__ inline Dual32 tsqrt(Dual32 a)
{
	const float s = __builtin_amdgcn_sqrtf(a.v);
	return Dual32(s, a.d * (0.5f * __builtin_amdgcn_rcpf(s)));
}
__device__ inline double tpow(double a, double b) { return pow(a, b); }
__device__ inline float tpow(float a, double b) { return powf(a, (float)b); }
__device__ inline Dual tpow(Dual a, double b)
{
	const double pv = pow(a.v, b);
	return Dual(pv, b * pow(a.v, b - 1.0) * a.d);
}
__device__ inline Dual32 tpow(Dual32 a, double b)
{
	const float pv = powf(a.v, (float)b);
	return Dual32(pv, (float)b * powf(a.v, (float)(b - 1.0)) * a.d);
}
__device__ inline double trcp(double a) { return 1.0 / a; }
__device__ inline float trcp(float a) { return __builtin_amdgcn_rcpf(a); }
__device__ inline Dual trcp(Dual a)
{
	const double r = 1.0 / a.v;
	return Dual(r, -a.d * r * r);
}
__device__ inline Dual32 trcp(Dual32 a)
{
	const float r = __builtin_amdgcn_rcpf(a.v);
	return Dual32(r, -a.d * r * r);
}
// drop the derivative part (neighbour traces held fixed for diagonal blocks)
__device__ inline double strip(double a) { return a; }
__device__ inline float strip(float a) { return a; }
__device__ inline Dual strip(Dual a) { return Dual(a.v, 0.0); }
__device__ inline Dual32 strip(Dual32 a) { return Dual32(a.v, 0.0f); }

template <typename T> struct RealOf { using type = T; };
template <> struct RealOf<Dual> { using type = double; };
template <> struct RealOf<Dual32> { using type = double; };   // FP64 vectors in and out

constexpr __host__ __device__ int ipow(int b, int e) { return e == 0 ? 1 : b * ipow(b, e - 1); }

// ------------------------------------------------------------------------------------------ fluxes
// f[k + D*eq] (flux_T.h:18-19 layout (node, dim, eq))
template <typename T, int D>
__device__ inline void euler_flux(const T* s, T* f)
{
	const T rho = s[0], E = s[D + 1];
	const T rho_inv = trcp(rho);
	T uvw[D];
#pragma unroll
	for (int k = 0; k < D; ++k) uvw[k] = rho_inv * s[1 + k];
	T V2 = uvw[0] * uvw[0];
#pragma unroll
	for (int k = 1; k < D; ++k) V2 += uvw[k] * uvw[k];
	const T p = T(kGM1) * (E - T(0.5) * rho * V2);
#pragma unroll
	for (int k = 0; k < D; ++k) f[k] = s[1 + k];
#pragma unroll
	for (int i = 0; i < D; ++i)
#pragma unroll
		for (int k = 0; k < D; ++k) f[k + D * (1 + i)] = (i == k) ? s[1 + i] * uvw[k] + p : s[1 + i] * uvw[k];
#pragma unroll
	for (int k = 0; k < D; ++k) f[k + D * (D + 1)] = (E + p) * uvw[k];
}

// n . F(W) for a state with precomputed primitives: [rho Vn, rho u_k Vn + p n_k, (E+p) Vn]
template <typename T, int D>
__device__ inline void euler_normal_flux(const T* W, const T* uvw, T p, const T* n, T* Fn)
{
	T Vn = uvw[0] * n[0];
#pragma unroll
	for (int k = 1; k < D; ++k) Vn += uvw[k] * n[k];
	Fn[0] = W[0] * Vn;
#pragma unroll
	for (int k = 0; k < D; ++k) Fn[1 + k] = W[1 + k] * Vn + p * n[k];
	Fn[D + 1] = (W[D + 1] + p) * Vn;
}

// Lax-Friedrichs: 0.5 [n.(F_L + F_R) + lambda (W_L - W_R)], lambda = max(|V| + c) over the sides
template <typename T, int D>
__device__ inline void numflux_lf(const T* WL, const T* WR, const T* n, T* nnf)
{
	constexpr int NV = D + 2;
	T uL[D], uR[D];
	const T riL = trcp(WL[0]), riR = trcp(WR[0]);
	T V2L = T(0.0), V2R = T(0.0);
#pragma unroll
	for (int k = 0; k < D; ++k) {
		uL[k] = WL[1 + k] * riL;
		uR[k] = WR[1 + k] * riR;
		V2L += uL[k] * uL[k];
		V2R += uR[k] * uR[k];
	}
	const T pL = T(kGM1) * (WL[D + 1] - T(0.5) * WL[0] * V2L);
	const T pR = T(kGM1) * (WR[D + 1] - T(0.5) * WR[0] * V2R);
	const T maxlL = tsqrt(V2L) + tsqrt(T(kGamma) * pL * riL);
	const T maxlR = tsqrt(V2R) + tsqrt(T(kGamma) * pR * riR);
	const T maxV = (rv(maxlL) > rv(maxlR)) ? maxlL : maxlR;
	T FL[NV], FR[NV];
	euler_normal_flux<T, D>(WL, uL, pL, n, FL);
	euler_normal_flux<T, D>(WR, uR, pR, n, FR);
#pragma unroll
	for (int eq = 0; eq < NV; ++eq) nnf[eq] = T(0.5) * ((FL[eq] + FR[eq]) + maxV * (WL[eq] - WR[eq]));
}

template <typename T>
__device__ inline T min_abs_real(T a, T b)
{
	T c = (fabs(rv(a)) < fabs(rv(b))) ? a : b;
	if (rv(c) < 0) c = T(-1.0) * c;
	return c;
}
template <typename T>
__device__ inline T max_abs_real(T a, T b)
{
	T c = (fabs(rv(a)) > fabs(rv(b))) ? a : b;
	if (rv(c) < 0) c = T(-1.0) * c;
	return c;
}

template <typename T, int D>
__device__ inline void numflux_roe(const T* WL, const T* WR, const T* n, T* nnf)
{
	const T n1 = n[0], n2 = (D > 1) ? n[1] : T(0.0), n3 = (D > 2) ? n[2] : T(0.0);
	const T rhoL = WL[0], rhouL = WL[1], rhovL = (D > 1) ? WL[2] : T(0.0), rhowL = (D > 2) ? WL[3] : T(0.0);
	const T EL = WL[D + 1];
	const T rhoR = WR[0], rhouR = WR[1], rhovR = (D > 1) ? WR[2] : T(0.0), rhowR = (D > 2) ? WR[3] : T(0.0);
	const T ER = WR[D + 1];
	const T rhoL_inv = trcp(rhoL), rhoR_inv = trcp(rhoR);
	const T uL = rhouL * rhoL_inv, vL = rhovL * rhoL_inv, wL = rhowL * rhoL_inv;
	const T uR = rhouR * rhoR_inv, vR = rhovR * rhoR_inv, wR = rhowR * rhoR_inv;
	T V2L = uL * uL, V2R = uR * uR, VnL = n1 * uL, VnR = n1 * uR;
	if (D > 1) { V2L += vL * vL; V2R += vR * vR; VnL += n2 * vL; VnR += n2 * vR; }
	if (D > 2) { V2L += wL * wL; V2R += wR * wR; VnL += n3 * wL; VnR += n3 * wR; }
	const T pL = T(kGM1) * (EL - T(0.5) * rhoL * V2L), pR = T(kGM1) * (ER - T(0.5) * rhoR * V2R);
	const T HL = (EL + pL) * rhoL_inv, HR = (ER + pR) * rhoR_inv;

	const T r = tsqrt(rhoR * rhoL_inv), rP1i = trcp(r + T(1.0));
	const T rho = r * rhoL;
	const T u = (r * uR + uL) * rP1i, v = (r * vR + vL) * rP1i, w = (r * wR + wL) * rP1i;
	const T H = (r * HR + HL) * rP1i;
	T Vn = n1 * u, V2 = u * u;
	if (D > 1) { Vn += n2 * v; V2 += v * v; }
	if (D > 2) { Vn += n3 * w; V2 += w * w; }
	const T c2 = T(kGM1) * (H - T(0.5) * V2), c = tsqrt(c2);

	const T l1 = min_abs_real(VnL - c, Vn - c);
	const T l5 = max_abs_real(VnR + c, Vn + c);
	const T l234 = (rv(Vn) > 0) ? Vn : T(-1.0) * Vn;

	const T drho = rhoR - rhoL;
	const T drhou = rhoR * uR - rhoL * uL, drhov = rhoR * vR - rhoL * vL, drhow = rhoR * wR - rhoL * wL;
	const T dE = ER - EL, dp = pR - pL, dVn = VnR - VnL;
	const T lc1 = T(0.5) * (l5 + l1) - l234, lc2 = T(0.5) * (l5 - l1);
	const T ci = trcp(c);
	const T disInter1 = lc1 * dp * (ci * ci) + lc2 * rho * dVn * ci;
	const T disInter2 = lc1 * rho * dVn + lc2 * dp * ci;

	const T rhoVnL = rhoL * VnL, rhoVnR = rhoR * VnR, pLR = pL + pR;
	int k = 0;
	nnf[k++] = T(0.5) * ((rhoVnL + rhoVnR) - (l234 * drho + disInter1));
	nnf[k++] = T(0.5) * ((rhoVnL * uL + rhoVnR * uR + n1 * pLR) - (l234 * drhou + disInter1 * u + disInter2 * n1));
	if (D > 1)
		nnf[k++] = T(0.5) * ((rhoVnL * vL + rhoVnR * vR + n2 * pLR) - (l234 * drhov + disInter1 * v + disInter2 * n2));
	if (D > 2)
		nnf[k++] = T(0.5) * ((rhoVnL * wL + rhoVnR * wR + n3 * pLR) - (l234 * drhow + disInter1 * w + disInter2 * n3));
	nnf[k++] = T(0.5) * ((VnL * (EL + pL) + VnR * (ER + pR)) - (l234 * dE + disInter1 * H + disInter2 * Vn));
}

// boundary state, n is the unit normal (primitive free stream in bc_prim)
template <typename T, int D>
__device__ inline void bc_state(int bc, const T* sL, const T* n, const double* bc_prim, T* sR)
{
	// results in scalars first, one straight-line store sequence at the end (no dynamic indexing)
	T o_rho, o_E, o_m[D];
	if (bc == 2) {  // slip wall
		T rVn = n[0] * sL[1];
#pragma unroll
		for (int k = 1; k < D; ++k) rVn += n[k] * sL[1 + k];
		o_rho = sL[0];
		o_E = sL[D + 1];
#pragma unroll
		for (int k = 0; k < D; ++k) o_m[k] = sL[1 + k] - T(2.0) * rVn * n[k];
	} else {
		const T rho_l = sL[0];
		T uvw_l[D];
		T V2l = T(0.0);
#pragma unroll
		for (int k = 0; k < D; ++k) { uvw_l[k] = sL[1 + k] / rho_l; V2l += uvw_l[k] * uvw_l[k]; }
		const T p_l = T(kGM1) * (sL[D + 1] - T(0.5) * rho_l * V2l);
		const T rho_r = T(bc_prim[0]), p_r = T(bc_prim[D + 1]);
		T uvw_r[D];
#pragma unroll
		for (int k = 0; k < D; ++k) uvw_r[k] = T(bc_prim[1 + k]);
		const T c_l = tsqrt(T(kGamma) * p_l / rho_l);
		const T c_r = tsqrt(T(kGamma) * p_r / rho_r);
		T Vn_l = T(0.0), Vn_r = T(0.0);
#pragma unroll
		for (int k = 0; k < D; ++k) { Vn_l += n[k] * uvw_l[k]; Vn_r += n[k] * uvw_r[k]; }
		const T R_l = Vn_l + T(2.0 / kGM1) * c_l;
		const T R_r = Vn_r - T(2.0 / kGM1) * c_r;
		const T Vn = T(0.5) * (R_l + R_r), cc = T(0.25 * kGM1) * (R_l - R_r);
		T rho, p, uvw[D];
		const bool super = fabs(rv(Vn)) >= fabs(rv(cc));
		const bool inlet = rv(Vn) < 0.0;
		if (super) {
			rho = inlet ? rho_r : rho_l;
			p = inlet ? p_r : p_l;
#pragma unroll
			for (int k = 0; k < D; ++k) uvw[k] = inlet ? uvw_r[k] : uvw_l[k];
		} else {
			const T s_b = inlet ? p_r / tpow(rho_r, kGamma) : p_l / tpow(rho_l, kGamma);
			rho = tpow(T(1.0 / kGamma) * cc * cc / s_b, 1.0 / kGM1);
#pragma unroll
			for (int k = 0; k < D; ++k) {
				const T ut = inlet ? uvw_r[k] - Vn_r * n[k] : uvw_l[k] - Vn_l * n[k];
				uvw[k] = Vn * n[k] + ut;
			}
			p = T(1.0 / kGamma) * cc * cc * rho;
		}
		T V2 = uvw[0] * uvw[0];
#pragma unroll
		for (int k = 1; k < D; ++k) V2 += uvw[k] * uvw[k];
		o_rho = rho;
#pragma unroll
		for (int k = 0; k < D; ++k) o_m[k] = rho * uvw[k];
		o_E = p / T(kGM1) + T(0.5) * rho * V2;
	}
	sR[0] = o_rho;
#pragma unroll
	for (int k = 0; k < D; ++k) sR[1 + k] = o_m[k];
	sR[D + 1] = o_E;
}

// contravariant Euler fluxes fr[d0] = sum_k m[k + D d0] F_k at one node, weighted by wq
template <typename T, typename R, int D>
__device__ inline void euler_contravariant_flux(const T* s, const R* m, R wq, T (*fr)[D + 2])
{
	const T ri = trcp(s[0]);
	T uvw[D];
	T V2 = T(0.0);
#pragma unroll
	for (int k = 0; k < D; ++k) { uvw[k] = s[1 + k] * ri; V2 += uvw[k] * uvw[k]; }
	const T p = T(kGM1) * (s[D + 1] - T(0.5) * s[0] * V2);
	const T Ep = s[D + 1] + p;
#pragma unroll
	for (int d0 = 0; d0 < D; ++d0) {
		T Vm = m[0 + D * d0] * uvw[0];
#pragma unroll
		for (int k = 1; k < D; ++k) Vm += m[k + D * d0] * uvw[k];
		fr[d0][0] = wq * (s[0] * Vm);
#pragma unroll
		for (int k = 0; k < D; ++k) fr[d0][1 + k] = wq * (s[1 + k] * Vm + p * m[k + D * d0]);
		fr[d0][D + 1] = wq * (Ep * Vm);
	}
}

// negated viscous flux fv[k + 2 * eq] of state s and gradient g[v + 4 * k] (2-D): tau with the -2/3 mu div V
// part (compute_tau, flux_navier_stokes_T.c:778-799), scaled temperature gradient d(E/rho) - d(V^2)/2
// (compute_dTs :904-926), energy flux gamma/Pr mu dTs + u . tau; constant viscosity
template <typename S>
__device__ inline void ns_viscous_flux2(const S* s, const S* g, S mu, S Pr, S* fv)
{
	const S rho = s[0], E = s[3], ri = trcp(rho);
	const S u[2] = { s[1] * ri, s[2] * ri };
	S du[2][2], tau[2][2], dTs[2];
#pragma unroll
	for (int k = 0; k < 2; ++k)
#pragma unroll
		for (int j = 0; j < 2; ++j) du[k][j] = ri * (g[(1 + k) + 4 * j] - u[k] * g[0 + 4 * j]);
	const S divV = du[0][0] + du[1][1];
#pragma unroll
	for (int i = 0; i < 2; ++i)
#pragma unroll
		for (int j = 0; j < 2; ++j) tau[i][j] = mu * (du[i][j] + du[j][i]) - (i == j ? S(2.0 / 3.0) * mu * divV : S(0.0));
#pragma unroll
	for (int j = 0; j < 2; ++j) {
		const S dV2 = S(2.0) * (u[0] * du[0][j] + u[1] * du[1][j]);
		dTs[j] = ri * ri * (g[3 + 4 * j] * rho - E * g[0 + 4 * j]) - S(0.5) * dV2;
	}
#pragma unroll
	for (int j = 0; j < 2; ++j) {
		fv[j + 2 * 0] = S(0.0);
		fv[j + 2 * 1] = -tau[0][j];
		fv[j + 2 * 2] = -tau[1][j];
		fv[j + 2 * 3] = -(S(kGamma) / Pr * mu * dTs[j] + u[0] * tau[j][0] + u[1] * tau[j][1]);
	}
}

// Viscosity at a state: constant mu, or Sutherland's formula mu_s = 8e-2 mu sqrt(T) / (1 + 112 / T) with
// T = p / (rho r_s) (compute_mu_sutherland_T, solution_navier_stokes_T.c:246-255; C1_SUTHERLAND, C2_SUTHERLAND,
// solution_navier_stokes.h:30-33). A compile-time switch: a runtime branch here doubled the residual kernels'
// VGPRs and cost the constant-viscosity kernels a third of their occupancy.
template <bool SUTHERLAND, typename S>
__device__ inline S ns_mu(const S* s, S mu, double r_s)
{
	if constexpr (!SUTHERLAND) return mu;
	const S ri = trcp(s[0]);
	const S V2 = (s[1] * s[1] + s[2] * s[2]) * ri * ri;
	const S p = S(kGM1) * (s[3] - S(0.5) * s[0] * V2);
	const S T = p * ri * S(1.0 / r_s);
	return S(8e-2) * mu * tsqrt(T) * trcp(S(1.0) + S(112.0) * trcp(T));
}

// Navier-Stokes wall data: the provided normal viscous energy flux of the no-slip flux walls, and the rotating
// wall's angular velocity, density and total energy (dgx_config wall_*)
struct NsWalls {
	double nfE, omega, rho, E;
};

// Boundary (ghost) state of a 2-D Navier-Stokes boundary face node: Riemann (boundary_euler_T.c:99-180); the
// no-slip adiabatic / diabatic walls keep rho and E and negate the momentum (-rho u_l + 2 rho u_ex, u_ex = 0;
// constructor_Boundary_Value_T_navier_stokes_no_slip_flux_general, boundary_navier_stokes_T.c:260-290); the
// rotating wall is 2 s_ex - s_l with s_ex = (rho_b, rho_b u_ex, E_b), u_ex = omega (-y, x) at the node xy
// (no_slip_all_general :200-240, compute_uvw_ex_rotating :505-520).  The boundary gradient is the interior one.
template <typename T>
__device__ inline void ns_wall_state(int bc, const T* sL, const T* n, const double* bc_prim, const NsWalls& w,
                                     const double* xy, T* sB)
{
	if (bc == DGX_BC_RIEMANN) {
		bc_state<T, 2>(bc, sL, n, bc_prim, sB);
		return;
	}
	if (bc == DGX_BC_NOSLIP_ROTATING) {
		const double ux = -w.omega * xy[1], uy = w.omega * xy[0];
		sB[0] = -sL[0] + T(2.0 * w.rho);
		sB[1] = -sL[1] + T(2.0 * w.rho * ux);
		sB[2] = -sL[2] + T(2.0 * w.rho * uy);
		sB[3] = -sL[3] + T(2.0 * w.E);
		return;
	}
	sB[0] = sL[0];
	sB[1] = -sL[1];
	sB[2] = -sL[2];
	sB[3] = sL[3];
}

}  // namespace dgx
