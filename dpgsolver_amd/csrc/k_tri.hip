// k_tri.hip -- TRI elements of 2-D mixed meshes and the faces that touch them (FP64, Euler and BR2 Navier-Stokes).
//
// TRI group kernels.  A TRI's operators are dense (every WSH node contributes to every cubature and face point),
// so a group kernel is a batch of small dense products: per element s = I u at the WV cubature, the
// contravariant fluxes, R = sum_d tw1_d fr_d - sum_f tw0_f z_f (compute_volume_rlhs_T.c:145-177,
// compute_face_rlhs_T.c:297-359).  One workgroup of 256 lanes takes a batch of EPB elements (as many as give
// every lane a cubature point, within an LDS budget): the batch's state and face blocks are read contiguously
// into LDS, each stage maps the lanes over the (element, point) pairs of the whole batch, and the order's
// operators are staged in LDS once per workgroup over a strided batch loop.
//   k_tri_trace: the state's traces at each face's own GL points (Euler, before the face kernels);
//   k_tri_rhs:   the Euler residual from the lifted face fluxes z;
//   k_tri_gv:    (BR2) the volume part of the weak gradient G_v = M^-1 I^T W (J grad s)
//                (compute_grad_coef_dg_T.c:171-307), its face traces, and the state traces (affine FP64 groups:
//                k_tri_gv_mfma on the matrix cores);
//   k_tri_ns:    G = G_v + sum_f M^-1 tw0_f y_f, the volume flux F_euler + fv (flux_navier_stokes_T.c:1005-1040)
//                and the residual.
// M^-1 and the lift traces K_f = Ifc_f M^-1 tw0_f: the reference ones divided by det J on an affine group, each
// element's own otherwise.
//
// Faces with a TRI side (k_mixf, k_mixf_ns): 2, 4 or 8 lanes per face (the next power of two >= Nf), four
// wavefronts per workgroup sharing one copy of the operators in LDS, each face's LDS steps synchronised within
// its wavefront; one launch per face-point count Nf = max(p_L, p_R) + 1 (compute_face_reference_order,
// face_solver.c:123-131; the faces arrive sorted by (Nf, types, orders), so a wavefront rarely mixes shapes).
// The Navier-Stokes kernel forms the two sides' face gradients one after the other in the same LDS block, which
// keeps its LDS per face at 16 Nf doubles.  Each side's own values
// (TRI: its traces; QUAD: its GLL face nodes) go through LDS and are interpolated to the face's GL points (the
// right side in the left ordering); the numerical flux uses n J_f of the lower-order side
// (get_dominant_geom_vol_side_index, face_solver.c:95-107); each side's z = P^T (w J_f F) / w_own at its own
// GL points (the right side negated, own ordering) is lifted exactly by its own test functions.  BR2: both
// sides use q = n J_f (s_R - s_L) / 2, each side's projection y = P^T w q / w_own (written after z for the
// residual kernels), its face gradient h = trace G_v + eta (lift trace) y with eta = 1.01 NFMAX = 4.04
// (compute_face_rlhs_dg_T.c:243-300; TRI: K_f y, QUAD: kappa y), and the central n . (fv_L + fv_R) / 2
// (numerical_flux_central_T.c:76-104) added to the inviscid flux.
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "k_residual.h"

namespace dgx {

namespace {

constexpr int kNV = 4;
constexpr int kTriLds = 6144;     // LDS budget of one TRI workgroup, in doubles (48 KB)
constexpr int kTriLdsNs = 8192;   // the Navier-Stokes kernels (64 KB; their registers allow two per CU anyway)

__device__ inline int quad_face_node(int N, int lf, int i)
{
	const int fd = lf >> 1, fix = (lf & 1) ? N - 1 : 0;
	return fd == 0 ? fix + N * i : i + N * fix;
}

constexpr int tri_np(int P) { return (P + 1) * (P + 2) / 2; }
constexpr int tri_nc(int P) { return P == 1 ? 3 : P == 2 ? 6 : P == 3 ? 12 : P == 4 ? 16 : 25; }   // WV of degree 2p

// elements per workgroup: enough to give every lane a point, within the LDS budget
constexpr int tri_epb(int lanes_per_elem, int ops, int per_elem, int budget = kTriLds)
{
	const int fill = 256 / lanes_per_elem > 0 ? 256 / lanes_per_elem : 1;
	const int fit = (budget - ops) / per_elem;
	return fill < fit ? fill : (fit > 0 ? fit : 1);
}

// each kernel's batch size (LDS: operators, then per element)
constexpr int epb_trace(int P) { return tri_epb(3 * (P + 1), 3 * (P + 1) * tri_np(P), kNV * tri_np(P)); }
constexpr int epb_rhs(int P)
{
	return tri_epb(tri_nc(P), 3 * tri_nc(P) * tri_np(P) + 3 * tri_np(P) * (P + 1),
	               kNV * tri_np(P) + 2 * kNV * tri_nc(P) + 3 * kNV * (P + 1) + 4 * tri_nc(P));
}
constexpr int epb_gv(int P)
{
	return tri_epb(tri_nc(P), 3 * tri_nc(P) * tri_np(P) + 3 * (P + 1) * tri_np(P) + tri_nc(P) + tri_np(P) * tri_np(P),
	               kNV * tri_np(P) + 2 * kNV * tri_nc(P) + 2 * kNV * tri_np(P) + 4 * tri_nc(P), kTriLdsNs);
}
constexpr int epb_ns(int P)
{
	return tri_epb(tri_nc(P), 3 * tri_nc(P) * tri_np(P) + 3 * tri_np(P) * (P + 1) + tri_np(P) * tri_np(P),
	               kNV * tri_np(P) + 9 * kNV * (P + 1) + 2 * kNV * tri_np(P) + 2 * kNV * tri_nc(P) + 4 * tri_nc(P),
	               kTriLdsNs);
}

template <typename T>
struct TriArgsT {
	const T* u;            // the group's state, E x nv x Np
	T* r;                  // the group's residual
	const T* met;          // E x 4 x Nc
	const T* z;            // the group's face blocks, E x 3 x zblk x N
	T* tr;                 // E x 3 x nv x N
	T* gv;                 // E x 2 x nv x Np
	T* gvt;                // E x 3 x 2 x nv x N
	const T* minv;         // curved groups: E x Np x Np
	const T* jinv;         // affine groups: E
	const TriOps<T>* ops;
	double mu, Pr;
	int E, affine;
};
using TriArgs = TriArgsT<double>;

template <typename T, int P>
__global__ void __launch_bounds__(256) k_tri_trace(TriArgsT<T> a)
{
	constexpr int Np = tri_np(P), N = P + 1, F3 = 3 * N, EPB = epb_trace(P);
	__shared__ T sIfc[F3 * Np];
	__shared__ T sU[EPB * kNV * Np];
	const int t = threadIdx.x;
	for (int i = t; i < F3 * Np; i += 256) sIfc[i] = a.ops->Ifc[i / (N * Np)][i % (N * Np)];
	for (int e0 = blockIdx.x * EPB; e0 < a.E; e0 += gridDim.x * EPB) {
		const int ne = min(EPB, a.E - e0);
		__syncthreads();
		for (int i = t; i < ne * kNV * Np; i += 256) sU[i] = a.u[(size_t)e0 * kNV * Np + i];
		__syncthreads();
		for (int i = t; i < ne * F3; i += 256) {
			const int sl = i / F3, k = i - sl * F3, lf = k / N, b = k - lf * N;
			const T* U = sU + sl * kNV * Np;
			T acc[kNV] = {};
#pragma unroll
			for (int m = 0; m < Np; ++m) {
				const T c = sIfc[k * Np + m];
#pragma unroll
				for (int v = 0; v < kNV; ++v) acc[v] += c * U[v * Np + m];
			}
			T* out = a.tr + ((size_t)(e0 + sl) * 3 + lf) * kNV * N + b;
#pragma unroll
			for (int v = 0; v < kNV; ++v) out[v * N] = acc[v];
		}
	}
}

template <typename T, int P>
__global__ void __launch_bounds__(256) k_tri_rhs(TriArgsT<T> a)
{
	constexpr int Np = tri_np(P), Nc = tri_nc(P), N = P + 1, ZB = kNV * N, EPB = epb_rhs(P);
	__shared__ T sI[Nc * Np], sDrW[Np * Nc], sDsW[Np * Nc], sTw0[3 * Np * N];
	__shared__ T sU[EPB * kNV * Np], sFR[EPB * 2 * kNV * Nc], sZ[EPB * 3 * ZB], sM[EPB * 4 * Nc];
	const int t = threadIdx.x;
	for (int i = t; i < Nc * Np; i += 256) {
		sI[i] = a.ops->I[i];
		sDrW[i] = a.ops->DrW[i];
		sDsW[i] = a.ops->DsW[i];
	}
	for (int i = t; i < 3 * Np * N; i += 256) sTw0[i] = a.ops->tw0[i / (Np * N)][i % (Np * N)];
	for (int e0 = blockIdx.x * EPB; e0 < a.E; e0 += gridDim.x * EPB) {
		const int ne = min(EPB, a.E - e0);
		__syncthreads();   // the operators are staged / the previous batch's buffers are read
		// every global read of the batch at once (one round trip), contiguous
		for (int i = t; i < ne * kNV * Np; i += 256) sU[i] = a.u[(size_t)e0 * kNV * Np + i];
		for (int i = t; i < ne * 3 * ZB; i += 256) sZ[i] = a.z[(size_t)e0 * 3 * ZB + i];
		for (int i = t; i < ne * 4 * Nc; i += 256) sM[i] = a.met[(size_t)e0 * 4 * Nc + i];
		__syncthreads();
		for (int i = t; i < ne * Nc; i += 256) {
			const int sl = i / Nc, q = i - sl * Nc;
			const T* U = sU + sl * kNV * Np;
			T s[kNV] = {};
#pragma unroll
			for (int m = 0; m < Np; ++m) {
				const T c = sI[q * Np + m];
#pragma unroll
				for (int v = 0; v < kNV; ++v) s[v] += c * U[v * Np + m];
			}
			const T* mp = sM + sl * 4 * Nc + q;
			const T m0 = mp[0], m1 = mp[Nc], m2 = mp[2 * Nc], m3 = mp[3 * Nc];
			T f[2 * kNV];
			euler_flux<T, 2>(s, f);
			T* FR = sFR + sl * 2 * kNV * Nc;
#pragma unroll
			for (int v = 0; v < kNV; ++v) {
				FR[v * Nc + q] = m0 * f[2 * v] + m1 * f[2 * v + 1];
				FR[(kNV + v) * Nc + q] = m2 * f[2 * v] + m3 * f[2 * v + 1];
			}
		}
		__syncthreads();
		for (int i = t; i < ne * Np; i += 256) {
			const int sl = i / Np, m = i - sl * Np;
			const T* FR = sFR + sl * 2 * kNV * Nc;
			const T* Z = sZ + sl * 3 * ZB;
			T acc[kNV] = {};
#pragma unroll
			for (int q = 0; q < Nc; ++q) {
				const T cr = sDrW[m * Nc + q], cs = sDsW[m * Nc + q];
#pragma unroll
				for (int v = 0; v < kNV; ++v) acc[v] += cr * FR[v * Nc + q] + cs * FR[(kNV + v) * Nc + q];
			}
#pragma unroll
			for (int lf = 0; lf < 3; ++lf)
#pragma unroll
				for (int b = 0; b < N; ++b) {
					const T c = sTw0[(lf * Np + m) * N + b];
#pragma unroll
					for (int v = 0; v < kNV; ++v) acc[v] -= c * Z[(lf * kNV + v) * N + b];
				}
			T* R = a.r + (size_t)(e0 + sl) * kNV * Np + m;
#pragma unroll
			for (int v = 0; v < kNV; ++v) R[v * Np] = acc[v];
		}
	}
}

// The same residual with the dense products on the matrix cores (v_mfma_f64_16x16x4f64, FP64 in and out).
// A wavefront takes four elements; the 16 rows of a tile are (variable v, element slot e) as i = e + 4 v, so that
// with the gfx950 layout (A: lane l holds A[l % 16][l / 16]; B: lane l holds B[l / 16][l % 16]; D: lane l,
// register r holds D[4 r + l / 16][l % 16]) each lane ends up with all four variables of element l / 16 at one
// cubature point or node l % 16 of the tile:
//   S^T (16 x Nc) = U^T (16 x Np) I^T                    the state at the cubature, then the fluxes in registers;
//   R^T (16 x Np) = FR_r^T DrW^T + FR_s^T DsW^T - Z^T tw0^T   the weak divergence and the face lifts,
// with the fluxes turned from the D layout to the A layout through LDS.  Operand B (the order's operators) is
// staged in LDS in fragment order (one double per lane per instruction); the state and the face blocks are read
// straight into A fragments.  Padding rows / columns / k-slots are zero, so they add nothing.
typedef double dgx_d4 __attribute__((ext_vector_type(4)));

template <int P>
struct TriMfma {
	static constexpr int Np = tri_np(P), Nc = tri_nc(P), N = P + 1;
	static constexpr int KS1 = (Np + 3) / 4, NT1 = (Nc + 15) / 16;              // S: k over nodes, tiles over points
	static constexpr int KS2 = (Nc + 3) / 4, QP = 4 * KS2, MT = (Np + 15) / 16;  // R: k over points, tiles over nodes
	static constexpr int KF = (3 * N + 3) / 4;                                  // R: k over the face points
};

template <int P>
__global__ void __launch_bounds__(256) k_tri_rhs_mfma(TriArgs a)
{
	using T = TriMfma<P>;
	constexpr int Np = T::Np, Nc = T::Nc, N = T::N, KS1 = T::KS1, NT1 = T::NT1, KS2 = T::KS2, QP = T::QP;
	constexpr int MT = T::MT, KF = T::KF;
	__shared__ double sB1[KS1 * NT1 * 64], sBr[KS2 * MT * 64], sBs[KS2 * MT * 64], sBf[KF * MT * 64];
	__shared__ double sFR[4][4 * 2 * kNV * QP];   // per wavefront: [e][d][v][q]
	const int t = threadIdx.x;
	for (int i = t; i < KS1 * NT1 * 64; i += 256) {
		const int ln = i & 63, f = i >> 6, kk = f / NT1, nt = f - kk * NT1;
		const int q = 16 * nt + (ln & 15), m = 4 * kk + (ln >> 4);
		sB1[i] = (q < Nc && m < Np) ? a.ops->I[q * Np + m] : 0.0;
	}
	for (int i = t; i < KS2 * MT * 64; i += 256) {
		const int ln = i & 63, f = i >> 6, kk = f / MT, mt = f - kk * MT;
		const int m = 16 * mt + (ln & 15), q = 4 * kk + (ln >> 4);
		const bool ok = m < Np && q < Nc;
		sBr[i] = ok ? a.ops->DrW[m * Nc + q] : 0.0;
		sBs[i] = ok ? a.ops->DsW[m * Nc + q] : 0.0;
	}
	for (int i = t; i < KF * MT * 64; i += 256) {
		const int ln = i & 63, f = i >> 6, kf = f / MT, mt = f - kf * MT;
		const int m = 16 * mt + (ln & 15), kap = 4 * kf + (ln >> 4);
		sBf[i] = (m < Np && kap < 3 * N) ? 0.0 - a.ops->tw0[kap / N][m * N + kap % N] : 0.0;
	}
	__syncthreads();
	const int wave = t >> 6, l = t & 63;
	const int ea = (l & 15) & 3, va = (l & 15) >> 2, ka = l >> 4;   // A fragment: row l % 16, k slot l / 16
	const int ed = l >> 4, jd = l & 15;                              // D fragment: element l / 16, column l % 16
	double* F = sFR[wave];
	const int ngroups = (a.E + 3) / 4;
	for (int g0 = blockIdx.x * 4; g0 < ngroups; g0 += gridDim.x * 4) {   // the same trip count for every wave
		const int grp = g0 + wave;
		const bool gvalid = grp < ngroups;
		const int e0 = grp * 4;
		const int eA = gvalid ? min(e0 + ea, a.E - 1) : 0;   // a padding slot reads a real element (never stored)
		const int eD = gvalid ? min(e0 + ed, a.E - 1) : 0;
		double ua[KS1], za[KF];
#pragma unroll
		for (int kk = 0; kk < KS1; ++kk) {
			const int m = 4 * kk + ka;
			ua[kk] = (gvalid && m < Np) ? a.u[(size_t)eA * kNV * Np + va * Np + m] : 0.0;
		}
#pragma unroll
		for (int kf = 0; kf < KF; ++kf) {
			const int kap = 4 * kf + ka;
			za[kf] = (gvalid && kap < 3 * N) ? a.z[(size_t)eA * 3 * kNV * N + ((kap / N) * kNV + va) * N + kap % N] : 0.0;
		}
		// the state at the cubature, the fluxes there in registers, then into LDS in the A layout
#pragma unroll
		for (int nt = 0; nt < NT1; ++nt) {
			dgx_d4 acc = { 0.0, 0.0, 0.0, 0.0 };
#pragma unroll
			for (int kk = 0; kk < KS1; ++kk)
				acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ua[kk], sB1[(kk * NT1 + nt) * 64 + l], acc, 0, 0, 0);
			const int q = 16 * nt + jd;
			if (q < QP) {
				double fr0[kNV] = {}, fr1[kNV] = {};
				if (gvalid && q < Nc) {
					const double s[kNV] = { acc[0], acc[1], acc[2], acc[3] };
					const double* mp = a.met + (size_t)eD * 4 * Nc + q;
					const double m0 = mp[0], m1 = mp[Nc], m2 = mp[2 * Nc], m3 = mp[3 * Nc];
					double f[2 * kNV];
					euler_flux<double, 2>(s, f);
#pragma unroll
					for (int v = 0; v < kNV; ++v) {
						fr0[v] = m0 * f[2 * v] + m1 * f[2 * v + 1];
						fr1[v] = m2 * f[2 * v] + m3 * f[2 * v + 1];
					}
				}
#pragma unroll
				for (int v = 0; v < kNV; ++v) {
					F[((ed * 2 + 0) * kNV + v) * QP + q] = fr0[v];
					F[((ed * 2 + 1) * kNV + v) * QP + q] = fr1[v];
				}
			}
		}
		__syncthreads();
		// the weak divergence and the face lifts, tile by tile of nodes
#pragma unroll
		for (int mt = 0; mt < MT; ++mt) {
			dgx_d4 acc = { 0.0, 0.0, 0.0, 0.0 };
#pragma unroll
			for (int kk = 0; kk < KS2; ++kk) {
				const int q = 4 * kk + ka;
				const double ar = F[((ea * 2 + 0) * kNV + va) * QP + q];
				const double as = F[((ea * 2 + 1) * kNV + va) * QP + q];
				acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, sBr[(kk * MT + mt) * 64 + l], acc, 0, 0, 0);
				acc = __builtin_amdgcn_mfma_f64_16x16x4f64(as, sBs[(kk * MT + mt) * 64 + l], acc, 0, 0, 0);
			}
#pragma unroll
			for (int kf = 0; kf < KF; ++kf)
				acc = __builtin_amdgcn_mfma_f64_16x16x4f64(za[kf], sBf[(kf * MT + mt) * 64 + l], acc, 0, 0, 0);
			const int m = 16 * mt + jd;
			if (gvalid && m < Np && e0 + ed < a.E) {
				double* R = a.r + (size_t)(e0 + ed) * kNV * Np + m;
#pragma unroll
				for (int v = 0; v < kNV; ++v) R[v * Np] = acc[v];
			}
		}
		__syncthreads();   // the fluxes are read before the next group overwrites them
	}
}

// The Navier-Stokes TRI residual on the matrix cores (affine groups).  With the tiles of k_tri_rhs_mfma, the state
// S^T = U^T I^T and each component d of the gradient at the cubature,
//   G_d^T = Gv_d^T I^T + (Y_d^T / det J) L^T,   L = I M^-1 tw0_f side by side over the three faces,
// are three (variable, element) tiles, so a lane holds the state and the whole gradient of one element at one
// cubature point: F_euler + fv is evaluated there, and the divergence and face lifts follow as for Euler.
template <int P>
__global__ void __launch_bounds__(256) k_tri_ns_mfma(TriArgs a)
{
	using T = TriMfma<P>;
	constexpr int Np = T::Np, Nc = T::Nc, N = T::N, KS1 = T::KS1, NT1 = T::NT1, KS2 = T::KS2, QP = T::QP;
	constexpr int MT = T::MT, KF = T::KF, ZB = 3 * kNV * N;
	__shared__ double sB1[KS1 * NT1 * 64], sBL[KF * NT1 * 64], sBr[KS2 * MT * 64], sBs[KS2 * MT * 64],
	    sBf[KF * MT * 64];
	__shared__ double sFR[4][4 * 2 * kNV * QP];
	const int t = threadIdx.x;
	for (int i = t; i < KS1 * NT1 * 64; i += 256) {
		const int ln = i & 63, f = i >> 6, kk = f / NT1, nt = f - kk * NT1;
		const int q = 16 * nt + (ln & 15), m = 4 * kk + (ln >> 4);
		sB1[i] = (q < Nc && m < Np) ? a.ops->I[q * Np + m] : 0.0;
	}
	for (int i = t; i < KF * NT1 * 64; i += 256) {
		const int ln = i & 63, f = i >> 6, kf = f / NT1, nt = f - kf * NT1;
		const int q = 16 * nt + (ln & 15), kap = 4 * kf + (ln >> 4);
		sBL[i] = (q < Nc && kap < 3 * N) ? a.ops->L[q * 3 * N + kap] : 0.0;
	}
	for (int i = t; i < KS2 * MT * 64; i += 256) {
		const int ln = i & 63, f = i >> 6, kk = f / MT, mt = f - kk * MT;
		const int m = 16 * mt + (ln & 15), q = 4 * kk + (ln >> 4);
		const bool ok = m < Np && q < Nc;
		sBr[i] = ok ? a.ops->DrW[m * Nc + q] : 0.0;
		sBs[i] = ok ? a.ops->DsW[m * Nc + q] : 0.0;
	}
	for (int i = t; i < KF * MT * 64; i += 256) {
		const int ln = i & 63, f = i >> 6, kf = f / MT, mt = f - kf * MT;
		const int m = 16 * mt + (ln & 15), kap = 4 * kf + (ln >> 4);
		sBf[i] = (m < Np && kap < 3 * N) ? 0.0 - a.ops->tw0[kap / N][m * N + kap % N] : 0.0;
	}
	__syncthreads();
	const double mu = a.mu, Pr = a.Pr;
	const int wave = t >> 6, l = t & 63;
	const int ea = (l & 15) & 3, va = (l & 15) >> 2, ka = l >> 4;
	const int ed = l >> 4, jd = l & 15;
	double* F = sFR[wave];
	const int ngroups = (a.E + 3) / 4;
	for (int g0 = blockIdx.x * 4; g0 < ngroups; g0 += gridDim.x * 4) {
		const int grp = g0 + wave;
		const bool gvalid = grp < ngroups;
		const int e0 = grp * 4;
		const int eA = gvalid ? min(e0 + ea, a.E - 1) : 0;
		const int eD = gvalid ? min(e0 + ed, a.E - 1) : 0;
		const double js = gvalid ? a.jinv[eA] : 0.0;
		double ua[KS1], g0a[KS1], g1a[KS1], y0a[KF], y1a[KF], za[KF];
#pragma unroll
		for (int kk = 0; kk < KS1; ++kk) {
			const int m = 4 * kk + ka;
			const bool ok = gvalid && m < Np;
			ua[kk] = ok ? a.u[(size_t)eA * kNV * Np + va * Np + m] : 0.0;
			g0a[kk] = ok ? a.gv[((size_t)eA * 2 * kNV + va) * Np + m] : 0.0;
			g1a[kk] = ok ? a.gv[((size_t)eA * 2 * kNV + kNV + va) * Np + m] : 0.0;
		}
#pragma unroll
		for (int kf = 0; kf < KF; ++kf) {
			const int kap = 4 * kf + ka;
			const bool ok = gvalid && kap < 3 * N;
			const double* zf = a.z + (size_t)eA * 3 * ZB + (kap / N) * ZB + kap % N;
			za[kf] = ok ? zf[va * N] : 0.0;
			y0a[kf] = ok ? js * zf[kNV * N + va * N] : 0.0;
			y1a[kf] = ok ? js * zf[kNV * N + (kNV + va) * N] : 0.0;
		}
#pragma unroll
		for (int nt = 0; nt < NT1; ++nt) {
			dgx_d4 as = { 0.0, 0.0, 0.0, 0.0 }, ag0 = as, ag1 = as;
#pragma unroll
			for (int kk = 0; kk < KS1; ++kk) {
				const double b = sB1[(kk * NT1 + nt) * 64 + l];
				as = __builtin_amdgcn_mfma_f64_16x16x4f64(ua[kk], b, as, 0, 0, 0);
				ag0 = __builtin_amdgcn_mfma_f64_16x16x4f64(g0a[kk], b, ag0, 0, 0, 0);
				ag1 = __builtin_amdgcn_mfma_f64_16x16x4f64(g1a[kk], b, ag1, 0, 0, 0);
			}
#pragma unroll
			for (int kf = 0; kf < KF; ++kf) {
				const double b = sBL[(kf * NT1 + nt) * 64 + l];
				ag0 = __builtin_amdgcn_mfma_f64_16x16x4f64(y0a[kf], b, ag0, 0, 0, 0);
				ag1 = __builtin_amdgcn_mfma_f64_16x16x4f64(y1a[kf], b, ag1, 0, 0, 0);
			}
			const int q = 16 * nt + jd;
			if (q < QP) {
				double fr0[kNV] = {}, fr1[kNV] = {};
				if (gvalid && q < Nc) {
					const double s[kNV] = { as[0], as[1], as[2], as[3] };
					const double gq[2 * kNV] = { ag0[0], ag0[1], ag0[2], ag0[3], ag1[0], ag1[1], ag1[2], ag1[3] };
					const double* mp = a.met + (size_t)eD * 4 * Nc + q;
					const double m0 = mp[0], m1 = mp[Nc], m2 = mp[2 * Nc], m3 = mp[3 * Nc];
					double f[2 * kNV], fv[2 * kNV];
					euler_flux<double, 2>(s, f);
					ns_viscous_flux2<double>(s, gq, mu, Pr, fv);
#pragma unroll
					for (int v = 0; v < kNV; ++v) {
						const double f0 = f[2 * v] + fv[2 * v], f1 = f[2 * v + 1] + fv[2 * v + 1];
						fr0[v] = m0 * f0 + m1 * f1;
						fr1[v] = m2 * f0 + m3 * f1;
					}
				}
#pragma unroll
				for (int v = 0; v < kNV; ++v) {
					F[((ed * 2 + 0) * kNV + v) * QP + q] = fr0[v];
					F[((ed * 2 + 1) * kNV + v) * QP + q] = fr1[v];
				}
			}
		}
		__syncthreads();
#pragma unroll
		for (int mt = 0; mt < MT; ++mt) {
			dgx_d4 acc = { 0.0, 0.0, 0.0, 0.0 };
#pragma unroll
			for (int kk = 0; kk < KS2; ++kk) {
				const int q = 4 * kk + ka;
				const double ar = F[((ea * 2 + 0) * kNV + va) * QP + q];
				const double as2 = F[((ea * 2 + 1) * kNV + va) * QP + q];
				acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, sBr[(kk * MT + mt) * 64 + l], acc, 0, 0, 0);
				acc = __builtin_amdgcn_mfma_f64_16x16x4f64(as2, sBs[(kk * MT + mt) * 64 + l], acc, 0, 0, 0);
			}
#pragma unroll
			for (int kf = 0; kf < KF; ++kf)
				acc = __builtin_amdgcn_mfma_f64_16x16x4f64(za[kf], sBf[(kf * MT + mt) * 64 + l], acc, 0, 0, 0);
			const int m = 16 * mt + jd;
			if (gvalid && m < Np && e0 + ed < a.E) {
				double* R = a.r + (size_t)(e0 + ed) * kNV * Np + m;
#pragma unroll
				for (int v = 0; v < kNV; ++v) R[v * Np] = acc[v];
			}
		}
		__syncthreads();
	}
}

// The BR2 volume gradient of an affine TRI group on the matrix cores (FP64).  With the (variable, element) tiles of
// k_tri_rhs_mfma:
//   SR^T, SS^T (16 x Nc) = U^T Dr^T, U^T Ds^T        the reference derivatives at the cubature; a lane then holds the
//       four variables of one element at one point and forms X_d = w (J grad s)_d with the cofactors;
//   G_d^T (16 x Np) = (X_d^T / det J) (M^-1 I^T)^T  the volume gradient G_v at the nodes (written out);
//   T^T (16 x 3N) = U^T Ifc^T, G_d^T Ifc^T             the state and gradient traces at the own GL face points,
// with X and then G turned from the D layout to the A layout through a per-wavefront LDS block.
template <int P>
__global__ void __launch_bounds__(256) k_tri_gv_mfma(TriArgs a)
{
	using T = TriMfma<P>;
	constexpr int Np = T::Np, Nc = T::Nc, N = T::N, KS1 = T::KS1, NT1 = T::NT1, KS2 = T::KS2, QP = T::QP;
	constexpr int MT = T::MT, F3 = 3 * N, FT = (F3 + 15) / 16, NPP = 4 * KS1, K8 = 2 * kNV;
	constexpr int XS = 4 * 2 * kNV * QP, GS = 4 * 2 * kNV * NPP, WS = XS > GS ? XS : GS;
	__shared__ double sBr[KS1 * NT1 * 64], sBs[KS1 * NT1 * 64], sBK[KS2 * MT * 64], sBF[KS1 * FT * 64];
	__shared__ double sWq[QP];
	__shared__ double sX[4][WS];   // per wavefront: X [e][d][v][q], then G [e][d][v][m]
	const int t = threadIdx.x;
	for (int i = t; i < KS1 * NT1 * 64; i += 256) {
		const int ln = i & 63, f = i >> 6, kk = f / NT1, nt = f - kk * NT1;
		const int q = 16 * nt + (ln & 15), m = 4 * kk + (ln >> 4);
		const bool ok = q < Nc && m < Np;
		sBr[i] = ok ? a.ops->Dr[q * Np + m] : 0.0;
		sBs[i] = ok ? a.ops->Ds[q * Np + m] : 0.0;
	}
	for (int i = t; i < KS2 * MT * 64; i += 256) {
		const int ln = i & 63, f = i >> 6, kk = f / MT, mt = f - kk * MT;
		const int m = 16 * mt + (ln & 15), q = 4 * kk + (ln >> 4);
		sBK[i] = (m < Np && q < Nc) ? a.ops->MIt[m * Nc + q] : 0.0;
	}
	for (int i = t; i < KS1 * FT * 64; i += 256) {
		const int ln = i & 63, f = i >> 6, kk = f / FT, ft = f - kk * FT;
		const int k0 = 16 * ft + (ln & 15), m = 4 * kk + (ln >> 4);
		sBF[i] = (k0 < F3 && m < Np) ? a.ops->Ifc[k0 / N][(k0 % N) * Np + m] : 0.0;
	}
	for (int i = t; i < QP; i += 256) sWq[i] = i < Nc ? a.ops->w[i] : 0.0;
	__syncthreads();
	const int wave = t >> 6, l = t & 63;
	const int ea = (l & 15) & 3, va = (l & 15) >> 2, ka = l >> 4;   // A fragment: row l % 16, k slot l / 16
	const int ed = l >> 4, jd = l & 15;                              // D fragment: element l / 16, column l % 16
	double* W = sX[wave];
	const int ngroups = (a.E + 3) / 4;
	for (int g0 = blockIdx.x * 4; g0 < ngroups; g0 += gridDim.x * 4) {   // the same trip count for every wave
		const int grp = g0 + wave;
		const bool gvalid = grp < ngroups;
		const int e0 = grp * 4;
		const int eA = gvalid ? min(e0 + ea, a.E - 1) : 0;   // a padding slot reads a real element (never stored)
		const int eD = gvalid ? min(e0 + ed, a.E - 1) : 0;
		const bool store = gvalid && e0 + ed < a.E;
		double ua[KS1];
#pragma unroll
		for (int kk = 0; kk < KS1; ++kk) {
			const int m = 4 * kk + ka;
			ua[kk] = (gvalid && m < Np) ? a.u[(size_t)eA * kNV * Np + va * Np + m] : 0.0;
		}
		// W J grad s at the cubature, into LDS in the A layout
#pragma unroll
		for (int nt = 0; nt < NT1; ++nt) {
			dgx_d4 ar = { 0.0, 0.0, 0.0, 0.0 }, as = ar;
#pragma unroll
			for (int kk = 0; kk < KS1; ++kk) {
				ar = __builtin_amdgcn_mfma_f64_16x16x4f64(ua[kk], sBr[(kk * NT1 + nt) * 64 + l], ar, 0, 0, 0);
				as = __builtin_amdgcn_mfma_f64_16x16x4f64(ua[kk], sBs[(kk * NT1 + nt) * 64 + l], as, 0, 0, 0);
			}
			const int q = 16 * nt + jd;
			if (q < QP) {
				double x0[kNV] = {}, x1[kNV] = {};
				if (gvalid && q < Nc) {
					const double* mp = a.met + (size_t)eD * 4 * Nc + q;
					const double m0 = mp[0], m1 = mp[Nc], m2 = mp[2 * Nc], m3 = mp[3 * Nc], w = sWq[q];
#pragma unroll
					for (int v = 0; v < kNV; ++v) {
						x0[v] = w * (m0 * ar[v] + m2 * as[v]);
						x1[v] = w * (m1 * ar[v] + m3 * as[v]);
					}
				}
#pragma unroll
				for (int v = 0; v < kNV; ++v) {
					W[((ed * 2 + 0) * kNV + v) * QP + q] = x0[v];
					W[((ed * 2 + 1) * kNV + v) * QP + q] = x1[v];
				}
			}
		}
		__syncthreads();
		// G_v = (M^-1 I^T X) / det J at the nodes
		dgx_d4 g[MT][2];
		const double js = gvalid ? a.jinv[eD] : 0.0;
#pragma unroll
		for (int mt = 0; mt < MT; ++mt) {
			dgx_d4 a0 = { 0.0, 0.0, 0.0, 0.0 }, a1 = a0;
#pragma unroll
			for (int kk = 0; kk < KS2; ++kk) {
				const int q = 4 * kk + ka;
				const double b = sBK[(kk * MT + mt) * 64 + l];
				a0 = __builtin_amdgcn_mfma_f64_16x16x4f64(W[((ea * 2 + 0) * kNV + va) * QP + q], b, a0, 0, 0, 0);
				a1 = __builtin_amdgcn_mfma_f64_16x16x4f64(W[((ea * 2 + 1) * kNV + va) * QP + q], b, a1, 0, 0, 0);
			}
#pragma unroll
			for (int v = 0; v < kNV; ++v) {
				g[mt][0][v] = a0[v] * js;
				g[mt][1][v] = a1[v] * js;
			}
		}
		__syncthreads();   // X is read by every lane of the wavefront before G takes its place
#pragma unroll
		for (int mt = 0; mt < MT; ++mt) {
			const int m = 16 * mt + jd;
			if (m < NPP) {
#pragma unroll
				for (int d = 0; d < 2; ++d)
#pragma unroll
					for (int v = 0; v < kNV; ++v) W[((ed * 2 + d) * kNV + v) * NPP + m] = m < Np ? g[mt][d][v] : 0.0;
			}
			if (store && m < Np) {
				double* G = a.gv + (size_t)(e0 + ed) * K8 * Np + m;
#pragma unroll
				for (int d = 0; d < 2; ++d)
#pragma unroll
					for (int v = 0; v < kNV; ++v) G[(d * kNV + v) * Np] = g[mt][d][v];
			}
		}
		__syncthreads();
		// the state and G_v at the own GL face points
#pragma unroll
		for (int ft = 0; ft < FT; ++ft) {
			dgx_d4 at = { 0.0, 0.0, 0.0, 0.0 }, ag0 = at, ag1 = at;
#pragma unroll
			for (int kk = 0; kk < KS1; ++kk) {
				const int m = 4 * kk + ka;
				const double b = sBF[(kk * FT + ft) * 64 + l];
				at = __builtin_amdgcn_mfma_f64_16x16x4f64(ua[kk], b, at, 0, 0, 0);
				ag0 = __builtin_amdgcn_mfma_f64_16x16x4f64(W[((ea * 2 + 0) * kNV + va) * NPP + m], b, ag0, 0, 0, 0);
				ag1 = __builtin_amdgcn_mfma_f64_16x16x4f64(W[((ea * 2 + 1) * kNV + va) * NPP + m], b, ag1, 0, 0, 0);
			}
			const int k0 = 16 * ft + jd;
			if (store && k0 < F3) {
				const int lf = k0 / N, b = k0 - lf * N;
				const size_t blk = (size_t)(e0 + ed) * 3 + lf;
#pragma unroll
				for (int v = 0; v < kNV; ++v) {
					a.tr[(blk * kNV + v) * N + b] = at[v];
					a.gvt[(blk * K8 + v) * N + b] = ag0[v];
					a.gvt[(blk * K8 + kNV + v) * N + b] = ag1[v];
				}
			}
		}
		__syncthreads();   // G is read before the next group's X overwrites it
	}
}

// M^-1 row m of element e (affine groups: the reference row in LDS and 1 / det J)
template <int Np, typename T>
__device__ inline void tri_minv_row(const TriArgsT<T>& a, const T* sMinv, int e, int m, const T*& row,
                                    T& scale)
{
	if (a.affine) {
		row = sMinv + m * Np;
		scale = a.jinv[e];
	} else {
		row = a.minv + (size_t)e * Np * Np + m * Np;
		scale = 1.0;
	}
}

template <typename T, int P>
__global__ void __launch_bounds__(256) k_tri_gv(TriArgsT<T> a)
{
	constexpr int Np = tri_np(P), Nc = tri_nc(P), N = P + 1, F3 = 3 * N, K8 = 2 * kNV, EPB = epb_gv(P);
	__shared__ T sI[Nc * Np], sDr[Nc * Np], sDs[Nc * Np], sIfc[F3 * Np], sW[Nc], sMinv[Np * Np];
	__shared__ T sU[EPB * kNV * Np], sX[EPB * K8 * Nc], sB[EPB * K8 * Np], sM[EPB * 4 * Nc];
	const int t = threadIdx.x;
	for (int i = t; i < Nc * Np; i += 256) {
		sI[i] = a.ops->I[i];
		sDr[i] = a.ops->Dr[i];
		sDs[i] = a.ops->Ds[i];
	}
	for (int i = t; i < F3 * Np; i += 256) sIfc[i] = a.ops->Ifc[i / (N * Np)][i % (N * Np)];
	for (int i = t; i < Nc; i += 256) sW[i] = a.ops->w[i];
	for (int i = t; i < Np * Np; i += 256) sMinv[i] = a.ops->Minv[i];
	for (int e0 = blockIdx.x * EPB; e0 < a.E; e0 += gridDim.x * EPB) {
		const int ne = min(EPB, a.E - e0);
		__syncthreads();
		for (int i = t; i < ne * kNV * Np; i += 256) sU[i] = a.u[(size_t)e0 * kNV * Np + i];
		for (int i = t; i < ne * 4 * Nc; i += 256) sM[i] = a.met[(size_t)e0 * 4 * Nc + i];
		__syncthreads();
		for (int i = t; i < ne * Nc; i += 256) {   // W J grad s at the cubature
			const int sl = i / Nc, q = i - sl * Nc;
			const T* U = sU + sl * kNV * Np;
			T sr[kNV] = {}, ss[kNV] = {};
#pragma unroll
			for (int m = 0; m < Np; ++m) {
				const T cr = sDr[q * Np + m], cs = sDs[q * Np + m];
#pragma unroll
				for (int v = 0; v < kNV; ++v) {
					sr[v] += cr * U[v * Np + m];
					ss[v] += cs * U[v * Np + m];
				}
			}
			const T* mp = sM + sl * 4 * Nc + q;
			const T m0 = mp[0], m1 = mp[Nc], m2 = mp[2 * Nc], m3 = mp[3 * Nc], w = sW[q];
			T* X = sX + sl * K8 * Nc;
#pragma unroll
			for (int v = 0; v < kNV; ++v) {
				X[v * Nc + q] = w * (m0 * sr[v] + m2 * ss[v]);
				X[(kNV + v) * Nc + q] = w * (m1 * sr[v] + m3 * ss[v]);
			}
		}
		__syncthreads();
		for (int i = t; i < ne * Np; i += 256) {   // I^T (W J grad s)
			const int sl = i / Np, m = i - sl * Np;
			const T* X = sX + sl * K8 * Nc;
			T acc[K8] = {};
#pragma unroll
			for (int q = 0; q < Nc; ++q) {
				const T c = sI[q * Np + m];
#pragma unroll
				for (int k = 0; k < K8; ++k) acc[k] += c * X[k * Nc + q];
			}
			T* B = sB + sl * K8 * Np;
#pragma unroll
			for (int k = 0; k < K8; ++k) B[k * Np + m] = acc[k];
		}
		__syncthreads();
		for (int i = t; i < ne * Np; i += 256) {   // G_v = M^-1 (...), kept in sX for the traces
			const int sl = i / Np, m = i - sl * Np, e = e0 + sl;
			const T* row;
			T sc;
			tri_minv_row<Np>(a, sMinv, e, m, row, sc);
			const T* B = sB + sl * K8 * Np;
			T g[K8] = {};
#pragma unroll
			for (int n = 0; n < Np; ++n) {
				const T c = row[n];
#pragma unroll
				for (int k = 0; k < K8; ++k) g[k] += c * B[k * Np + n];
			}
			T* G = sX + sl * K8 * Nc;
#pragma unroll
			for (int k = 0; k < K8; ++k) {
				g[k] *= sc;
				G[k * Np + m] = g[k];
				a.gv[((size_t)e * K8 + k) * Np + m] = g[k];
			}
		}
		__syncthreads();
		for (int i = t; i < ne * F3; i += 256) {   // traces of G_v and of the state at the own GL face points
			const int sl = i / F3, k0 = i - sl * F3, lf = k0 / N, b = k0 - lf * N;
			const T* G = sX + sl * K8 * Nc;
			const T* U = sU + sl * kNV * Np;
			T acc[K8] = {}, st[kNV] = {};
#pragma unroll
			for (int m = 0; m < Np; ++m) {
				const T c = sIfc[k0 * Np + m];
#pragma unroll
				for (int k = 0; k < K8; ++k) acc[k] += c * G[k * Np + m];
#pragma unroll
				for (int v = 0; v < kNV; ++v) st[v] += c * U[v * Np + m];
			}
			const size_t blk = (size_t)(e0 + sl) * 3 + lf;
#pragma unroll
			for (int k = 0; k < K8; ++k) a.gvt[(blk * K8 + k) * N + b] = acc[k];
#pragma unroll
			for (int v = 0; v < kNV; ++v) a.tr[(blk * kNV + v) * N + b] = st[v];
		}
	}
}

template <typename T, int P>
__global__ void __launch_bounds__(256) k_tri_ns(TriArgsT<T> a)
{
	constexpr int Np = tri_np(P), Nc = tri_nc(P), N = P + 1, K8 = 2 * kNV, ZB = 3 * kNV * N, EPB = epb_ns(P);
	__shared__ T sI[Nc * Np], sDrW[Np * Nc], sDsW[Np * Nc], sTw0[3 * Np * N], sMinv[Np * Np];
	__shared__ T sU[EPB * kNV * Np], sZ[EPB * 3 * ZB], sG[EPB * K8 * Np], sFR[EPB * K8 * Nc], sM[EPB * 4 * Nc];
	const int t = threadIdx.x;
	for (int i = t; i < Nc * Np; i += 256) {
		sI[i] = a.ops->I[i];
		sDrW[i] = a.ops->DrW[i];
		sDsW[i] = a.ops->DsW[i];
	}
	for (int i = t; i < 3 * Np * N; i += 256) sTw0[i] = a.ops->tw0[i / (Np * N)][i % (Np * N)];
	for (int i = t; i < Np * Np; i += 256) sMinv[i] = a.ops->Minv[i];
	const T mu = a.mu, Pr = a.Pr;
	for (int e0 = blockIdx.x * EPB; e0 < a.E; e0 += gridDim.x * EPB) {
		const int ne = min(EPB, a.E - e0);
		__syncthreads();
		// every global read of the batch at once (one round trip), contiguous; G starts as G_v
		for (int i = t; i < ne * kNV * Np; i += 256) sU[i] = a.u[(size_t)e0 * kNV * Np + i];
		for (int i = t; i < ne * 3 * ZB; i += 256) sZ[i] = a.z[(size_t)e0 * 3 * ZB + i];
		for (int i = t; i < ne * K8 * Np; i += 256) sG[i] = a.gv[(size_t)e0 * K8 * Np + i];
		for (int i = t; i < ne * 4 * Nc; i += 256) sM[i] = a.met[(size_t)e0 * 4 * Nc + i];
		__syncthreads();
		// the face lifts of the gradient: sum_f tw0_f y_f at the nodes (into sFR, free until the fluxes)
		for (int i = t; i < ne * Np; i += 256) {
			const int sl = i / Np, m = i - sl * Np;
			const T* Z = sZ + sl * 3 * ZB;
			T acc[K8] = {};
#pragma unroll
			for (int lf = 0; lf < 3; ++lf)
#pragma unroll
				for (int b = 0; b < N; ++b) {
					const T c = sTw0[(lf * Np + m) * N + b];
#pragma unroll
					for (int k = 0; k < K8; ++k) acc[k] += c * Z[lf * ZB + kNV * N + k * N + b];
				}
			T* Lt = sFR + sl * K8 * Nc;
#pragma unroll
			for (int k = 0; k < K8; ++k) Lt[k * Np + m] = acc[k];
		}
		__syncthreads();
		for (int i = t; i < ne * Np; i += 256) {   // G = G_v + M^-1 t
			const int sl = i / Np, m = i - sl * Np, e = e0 + sl;
			const T* row;
			T sc;
			tri_minv_row<Np>(a, sMinv, e, m, row, sc);
			const T* Lt = sFR + sl * K8 * Nc;
			T acc[K8] = {};
#pragma unroll
			for (int n = 0; n < Np; ++n) {
				const T c = row[n];
#pragma unroll
				for (int k = 0; k < K8; ++k) acc[k] += c * Lt[k * Np + n];
			}
			T* G = sG + sl * K8 * Np;
#pragma unroll
			for (int k = 0; k < K8; ++k) G[k * Np + m] += sc * acc[k];
		}
		__syncthreads();
		for (int i = t; i < ne * Nc; i += 256) {
			const int sl = i / Nc, q = i - sl * Nc;
			const T* U = sU + sl * kNV * Np;
			const T* G = sG + sl * K8 * Np;
			T s[kNV] = {}, gq[K8] = {};
#pragma unroll
			for (int m = 0; m < Np; ++m) {
				const T c = sI[q * Np + m];
#pragma unroll
				for (int v = 0; v < kNV; ++v) s[v] += c * U[v * Np + m];
#pragma unroll
				for (int k = 0; k < K8; ++k) gq[k] += c * G[k * Np + m];
			}
			const T* mp = sM + sl * 4 * Nc + q;
			const T m0 = mp[0], m1 = mp[Nc], m2 = mp[2 * Nc], m3 = mp[3 * Nc];
			T f[K8], fv[K8];
			euler_flux<T, 2>(s, f);
			ns_viscous_flux2<T>(s, gq, mu, Pr, fv);
			T* FR = sFR + sl * K8 * Nc;
#pragma unroll
			for (int v = 0; v < kNV; ++v) {
				const T f0 = f[2 * v] + fv[2 * v], f1 = f[2 * v + 1] + fv[2 * v + 1];
				FR[v * Nc + q] = m0 * f0 + m1 * f1;
				FR[(kNV + v) * Nc + q] = m2 * f0 + m3 * f1;
			}
		}
		__syncthreads();
		for (int i = t; i < ne * Np; i += 256) {
			const int sl = i / Np, m = i - sl * Np;
			const T* FR = sFR + sl * K8 * Nc;
			const T* Z = sZ + sl * 3 * ZB;
			T acc[kNV] = {};
#pragma unroll
			for (int q = 0; q < Nc; ++q) {
				const T cr = sDrW[m * Nc + q], cs = sDsW[m * Nc + q];
#pragma unroll
				for (int v = 0; v < kNV; ++v) acc[v] += cr * FR[v * Nc + q] + cs * FR[(kNV + v) * Nc + q];
			}
#pragma unroll
			for (int lf = 0; lf < 3; ++lf)
#pragma unroll
				for (int b = 0; b < N; ++b) {
					const T c = sTw0[(lf * Np + m) * N + b];
#pragma unroll
					for (int v = 0; v < kNV; ++v) acc[v] -= c * Z[lf * ZB + v * N + b];
				}
			T* R = a.r + (size_t)(e0 + sl) * kNV * Np + m;
#pragma unroll
			for (int v = 0; v < kNV; ++v) R[v * Np] = acc[v];
		}
	}
}

// r <- M^-1 r per element (scale_rhs_by_m_inv_std, solve_dg.c:265-272), and max |M^-1 r| into dmax (the bits of a
// non-negative double order as the double does; NaN propagates as the largest)
constexpr int epb_minv(int P) { return tri_epb(tri_np(P), tri_np(P) * tri_np(P), kNV * tri_np(P)); }

template <int P>
__global__ void __launch_bounds__(256) k_tri_minv(TriArgs a, unsigned long long* dmax)
{
	constexpr int Np = tri_np(P), EPB = epb_minv(P);
	__shared__ double sMinv[Np * Np];
	__shared__ double sR[EPB * kNV * Np];
	const int t = threadIdx.x;
	for (int i = t; i < Np * Np; i += 256) sMinv[i] = a.ops->Minv[i];
	double lmax = 0.0;
	for (int e0 = blockIdx.x * EPB; e0 < a.E; e0 += gridDim.x * EPB) {
		const int ne = min(EPB, a.E - e0);
		__syncthreads();
		for (int i = t; i < ne * kNV * Np; i += 256) sR[i] = a.r[(size_t)e0 * kNV * Np + i];
		__syncthreads();
		for (int i = t; i < ne * Np; i += 256) {
			const int sl = i / Np, m = i - sl * Np, e = e0 + sl;
			const double* row;
			double sc;
			tri_minv_row<Np>(a, sMinv, e, m, row, sc);
			const double* R = sR + sl * kNV * Np;
			double acc[kNV] = {};
#pragma unroll
			for (int n = 0; n < Np; ++n) {
				const double c = row[n];
#pragma unroll
				for (int v = 0; v < kNV; ++v) acc[v] += c * R[v * Np + n];
			}
			double* out = a.r + (size_t)e * kNV * Np + m;
#pragma unroll
			for (int v = 0; v < kNV; ++v) {
				const double x = sc * acc[v];
				out[v * Np] = x;
				const double ax = fabs(x);
				lmax = (ax > lmax || ax != ax) ? ax : lmax;
			}
		}
	}
	if (dmax) {
		for (int off = 32; off > 0; off >>= 1) {
			const double o = __shfl_down(lmax, off);
			lmax = (o > lmax || o != o) ? o : lmax;
		}
		if ((threadIdx.x & 63) == 0) atomicMax(dmax, (unsigned long long)__double_as_longlong(lmax));
	}
}

// ------------------------------------------------------------------------------------------------ faces
template <typename T>
struct MixArgsT {
	const MixFace* faces;              // the whole sorted list
	const T* nJ;                  // Fmix x kHpMaxN x 2
	const T* u;                   // the whole state
	T* z;                         // the whole lifted-flux buffer
	const HpOps<T>* hops;
	const TriOps<T>* tops;        // indexed by p
	const T* ttr[kTriMaxP];       // TRI groups: state traces
	const T* tgvt[kTriMaxP];      // TRI groups: traces of G_v
	const T* tK[kTriMaxP];        // curved TRI groups: each element's K_f (affine groups: null)
	const T* tjinv[kTriMaxP];     // affine TRI groups: 1 / det J
	const T* qgvt[kHpMaxN + 1];   // QUAD groups by N: traces of G_v and kappa
	const T* qkap[kHpMaxN + 1];
	double bc_prim[5];
	double mu, Pr;
	NsWalls walls;                     // Navier-Stokes wall data
	const double* wxy;                 // rotating walls: coordinates of each face's GL points (Fmix x kHpMaxN x 2)
	int start, count;
};

// value j of variable v of one side at its own face points (TRI: trace at its GL points; QUAD: GLL face node)
template <typename T>
__device__ inline T side_value(const MixArgsT<T>& a, int kind, int N, int lf, long long uoff, int el, int tg, int v,
                                    int j)
{
	if (kind) return a.ttr[tg][(((size_t)el * 3 + lf) * kNV + v) * N + j];
	return a.u[uoff + (long long)v * N * N + quad_face_node(N, lf, j)];
}

// lanes per face: the next power of two >= the face-point count; four wavefronts per workgroup share one copy
// of the operators, and the faces of a wavefront are synchronised within the wavefront only
constexpr int mix_lpf(int NF) { return NF <= 2 ? 2 : NF <= 4 ? 4 : 8; }
constexpr int kMixWaves = 4;
constexpr int mix_fpb(int NF) { return kMixWaves * (64 / mix_lpf(NF)); }

// orders the LDS writes and reads of one wavefront (every lane of a face is in the same wavefront)
__device__ inline void wave_sync()
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// the operators of one face-point count NF, for the own point counts N = 2..NF (index N - 2)
template <typename T, int NF>
struct MixOps {
	T Ig[NF - 1][NF * NF];   // own GLL points -> the NF face points ([q * N + i])
	T Pg[NF - 1][NF * NF];   // own GL points -> the NF face points
	T winv[NF - 1][NF];
	T w[NF];
};

template <int NF, typename T>
__device__ inline void stage_mix_ops(MixOps<T, NF>& o, const HpOps<T>* h)
{
	for (int i = threadIdx.x; i < (NF - 1) * NF * NF; i += blockDim.x) {
		const int n = i / (NF * NF), k = i - n * NF * NF;
		o.Ig[n][k] = h->Ig[n + 2][NF][k];
		o.Pg[n][k] = h->Pg[n + 2][NF][k];
	}
	for (int i = threadIdx.x; i < (NF - 1) * NF; i += blockDim.x) o.winv[i / NF][i % NF] = h->winv[i / NF + 2][i % NF];
	for (int i = threadIdx.x; i < NF; i += blockDim.x) o.w[i] = h->w[NF][i];
}

// one side's state at face point q from its own values S[v * NF + i] (TRI: from GL points; QUAD: from GLL)
template <int NF, typename T>
__device__ inline void side_at_point(const MixOps<T, NF>& o, int kind, int Ns, int q, const T* S, T* out)
{
	const T* op = kind ? o.Pg[Ns - 2] : o.Ig[Ns - 2];
#pragma unroll
	for (int v = 0; v < kNV; ++v) {
		T acc = 0.0;
#pragma unroll
		for (int i = 0; i < NF; ++i)
			if (i < Ns) acc += op[q * Ns + i] * S[v * NF + i];
		out[v] = acc;
	}
}

// z = P^T g / w_own at own point j of one side (g: [q][v] at the NF face points, in the left ordering)
template <int NF, typename T>
__device__ inline void lift_side(const MixOps<T, NF>& o, int Ns, int j, bool rev, bool neg, const T* g, T* z)
{
	const T* Pg = o.Pg[Ns - 2];
#pragma unroll
	for (int v = 0; v < kNV; ++v) {
		T acc = 0.0;
#pragma unroll
		for (int q = 0; q < NF; ++q) acc += Pg[(rev ? NF - 1 - q : q) * Ns + j] * g[q * kNV + v];
		z[v * Ns + j] = neg ? T(0) - acc * o.winv[Ns - 2][j] : acc * o.winv[Ns - 2][j];
	}
}

// one side's values at its own face points, loaded into registers by lane j (j < Ns)
template <typename T>
__device__ inline void load_own(const MixArgsT<T>& a, int kind, int Ns, int lf, long long uoff, int el, int tg, int j,
                                T* own)
{
#pragma unroll
	for (int v = 0; v < kNV; ++v) own[v] = side_value(a, kind, Ns, lf, uoff, el, tg, v, j);
}

template <typename T, int NF, int FLUX>
__global__ void __launch_bounds__(256) k_mixf(MixArgsT<T> a)
{
	constexpr int LPF = mix_lpf(NF), FPB = mix_fpb(NF);
	__shared__ MixOps<T, NF> o;
	__shared__ T sF[FPB][2 * kNV * NF];   // both sides' own values [side][v][i], then w J_f F [q][v]
	const int slot = threadIdx.x / LPF, j = threadIdx.x % LPF;
	const int fl = blockIdx.x * FPB + slot;
	const bool valid = fl < a.count;
	const int fi = a.start + (valid ? fl : 0);
	const MixFace f = a.faces[fi];
	const bool inter = f.bc == DGX_BC_INTERIOR;
	// every global load of this lane first (one round trip), then the operators
	T oL[kNV], oR[kNV], nj0 = 1.0, nj1 = 0.0;
	if (valid && j < f.NL) load_own(a, f.kL, f.NL, f.lfL, f.uL, f.eL, f.tL, j, oL);
	if (valid && inter && j < f.NR) load_own(a, f.kR, f.NR, f.lfR, f.uR, f.eR, f.tR, j, oR);
	if (valid && j < NF) {
		const T* nj = a.nJ + ((size_t)fi * kHpMaxN + j) * 2;
		nj0 = nj[0];
		nj1 = nj[1];
	}
	stage_mix_ops<NF>(o, a.hops);
	T* F = sF[slot];
	if (valid && j < f.NL)
#pragma unroll
		for (int v = 0; v < kNV; ++v) F[v * NF + j] = oL[v];
	if (valid && inter && j < f.NR)
#pragma unroll
		for (int v = 0; v < kNV; ++v) F[kNV * NF + v * NF + j] = oR[v];
	__syncthreads();   // the operators and every face's own values
	const int q = j;
	T g[kNV];
	if (valid && j < NF) {
		const T Jf = sqrt(nj0 * nj0 + nj1 * nj1);
		const T nu[2] = { nj0 / Jf, nj1 / Jf };
		T sL[kNV], sR[kNV], nnf[kNV];
		side_at_point<NF>(o, f.kL, f.NL, q, F, sL);
		if (inter) side_at_point<NF>(o, f.kR, f.NR, f.rev ? NF - 1 - q : q, F + kNV * NF, sR);
		else bc_state<T, 2>(f.bc, sL, nu, a.bc_prim, sR);
		if constexpr (FLUX == DGX_FLUX_LF) numflux_lf<T, 2>(sL, sR, nu, nnf);
		else numflux_roe<T, 2>(sL, sR, nu, nnf);
#pragma unroll
		for (int v = 0; v < kNV; ++v) g[v] = o.w[q] * Jf * nnf[v];
	}
	wave_sync();   // the own values are read
	if (valid && j < NF)
#pragma unroll
		for (int v = 0; v < kNV; ++v) F[q * kNV + v] = g[v];
	wave_sync();
	if (valid && j < f.NL) lift_side<NF>(o, f.NL, j, false, false, F, a.z + f.zL);
	if (valid && inter && j < f.NR) lift_side<NF>(o, f.NR, j, f.rev != 0, true, F, a.z + f.zR);
}

// what a side's face gradient at its own GL point j needs from memory: the trace of G_v there (t[k]) and its
// row of the lift-trace operator (TRI: K_f row j times sc; QUAD: kappa in kr[0])
template <int NF, typename T>
__device__ inline void load_side_grad(const MixArgsT<T>& a, int kind, int Ns, int lf, int el, int tg, int j, T* t,
                                      T* kr, T& sc)
{
	if (kind) {
		const T* tp = a.tgvt[tg] + ((size_t)el * 3 + lf) * 2 * kNV * Ns + j;
#pragma unroll
		for (int k = 0; k < 2 * kNV; ++k) t[k] = tp[k * Ns];
		const T* K;
		if (a.tK[tg]) {
			K = a.tK[tg] + ((size_t)el * 3 + lf) * Ns * Ns + j * Ns;
			sc = T(1);
		} else {
			K = a.tops[Ns - 1].K[lf] + j * Ns;
			sc = a.tjinv[tg][el];
		}
#pragma unroll
		for (int b = 0; b < NF; ++b) kr[b] = b < Ns ? K[b] : T(0);
		return;
	}
	const T* tp = a.qgvt[Ns] + ((size_t)el * 4 + lf) * 2 * kNV * Ns + j;
#pragma unroll
	for (int k = 0; k < 2 * kNV; ++k) t[k] = tp[k * Ns];
	kr[0] = a.qkap[Ns][((size_t)el * 4 + lf) * Ns + j];
	sc = T(1);
}

// the face gradient h = trace G_v + eta (lift trace) y at own point j, y = Y[k * NF + b]
template <int NF, typename T>
__device__ inline void side_face_gradient(int kind, int Ns, int j, const T* t, const T* kr, T sc,
                                          const T* Y, T* h)
{
	constexpr T eta = T(1.01 * 4.0);
	if (kind) {
#pragma unroll
		for (int k = 0; k < 2 * kNV; ++k) {
			T acc = 0.0;
#pragma unroll
			for (int b = 0; b < NF; ++b)
				if (b < Ns) acc += kr[b] * Y[k * NF + b];
			h[k] = t[k] + eta * (sc * acc);
		}
		return;
	}
#pragma unroll
	for (int k = 0; k < 2 * kNV; ++k) h[k] = t[k] + eta * kr[0] * Y[k * NF + j];
}

template <typename T, int NF, int FLUX>
__global__ void __launch_bounds__(256) k_mixf_ns(MixArgsT<T> a)
{
	constexpr int LPF = mix_lpf(NF), FPB = mix_fpb(NF), K8 = 2 * kNV;
	__shared__ MixOps<T, NF> o;
	__shared__ T sY[FPB][2][K8 * NF];   // both sides' own values [side][v][i]; then each side's y [k][i], then h
	__shared__ T sQ[FPB][NF * K8];      // w q at the face points [q][k]; then w J_f F [q][v]
	const int slot = threadIdx.x / LPF, j = threadIdx.x % LPF;
	const int fl = blockIdx.x * FPB + slot;
	const bool valid = fl < a.count;
	const int fi = a.start + (valid ? fl : 0);
	const MixFace f = a.faces[fi];
	// the own values and normals first, then the operators
	T oL[kNV], oR[kNV], nj0 = 1.0, nj1 = 0.0;
	T tL[K8], tR[K8], kL[NF], kR[NF], scL = 1.0, scR = 1.0;
	const bool inter = f.bc == DGX_BC_INTERIOR;   // a boundary face has a TRI left side only
	const bool ownL = valid && j < f.NL, ownR = valid && inter && j < f.NR, point = valid && j < NF;
	if (ownL) load_own(a, f.kL, f.NL, f.lfL, f.uL, f.eL, f.tL, j, oL);
	if (ownR) load_own(a, f.kR, f.NR, f.lfR, f.uR, f.eR, f.tR, j, oR);
	if (point) {
		const T* nj = a.nJ + ((size_t)fi * kHpMaxN + j) * 2;
		nj0 = nj[0];
		nj1 = nj[1];
	}
	stage_mix_ops<NF>(o, a.hops);
	T* YL = sY[slot][0];
	T* YR = sY[slot][1];
	T* Q = sQ[slot];
	if (ownL)
#pragma unroll
		for (int v = 0; v < kNV; ++v) YL[v * NF + j] = oL[v];
	if (ownR)
#pragma unroll
		for (int v = 0; v < kNV; ++v) YL[kNV * NF + v * NF + j] = oR[v];
	__syncthreads();   // the operators and every face's own values
	// the gradient traces and lift-trace rows, needed from the face-gradient step on
	if (ownL) load_side_grad<NF>(a, f.kL, f.NL, f.lfL, f.eL, f.tL, j, tL, kL, scL);
	if (ownR) load_side_grad<NF>(a, f.kR, f.NR, f.lfR, f.eR, f.tR, j, tR, kR, scR);
	T sL[kNV], sR[kNV];
	const int q = j, qr = f.rev ? NF - 1 - q : q;
	if (point) {
		side_at_point<NF>(o, f.kL, f.NL, q, YL, sL);
		if (inter) {
			side_at_point<NF>(o, f.kR, f.NR, qr, YL + kNV * NF, sR);
		} else {   // the boundary state (Riemann or a no-slip wall) at this point
			const T Jf = sqrt(nj0 * nj0 + nj1 * nj1);
			const T nu[2] = { nj0 / Jf, nj1 / Jf };
			double xy[2] = { 0.0, 0.0 };
			if (f.bc == DGX_BC_NOSLIP_ROTATING) {
				xy[0] = a.wxy[((size_t)fi * kHpMaxN + q) * 2];
				xy[1] = a.wxy[((size_t)fi * kHpMaxN + q) * 2 + 1];
			}
			ns_wall_state<T>(f.bc, sL, nu, a.bc_prim, a.walls, xy, sR);
		}
#pragma unroll
		for (int v = 0; v < kNV; ++v) {
			const T dl = T(0.5) * (sR[v] - sL[v]);
			Q[q * K8 + v] = o.w[q] * (nj0 * dl);
			Q[q * K8 + kNV + v] = o.w[q] * (nj1 * dl);
		}
	}
	wave_sync();   // the own values are read and every q is written
	// projections y to each side's own GL points (written for the residual kernels, after the lifted flux)
	if (ownL) {
		const T* Pg = o.Pg[f.NL - 2];
#pragma unroll
		for (int k = 0; k < K8; ++k) {
			T acc = 0.0;
#pragma unroll
			for (int qq = 0; qq < NF; ++qq) acc += Pg[qq * f.NL + j] * Q[qq * K8 + k];
			acc *= o.winv[f.NL - 2][j];
			YL[k * NF + j] = acc;
			a.z[f.zL + kNV * f.NL + k * f.NL + j] = acc;
		}
	}
	if (ownR) {
		const T* Pg = o.Pg[f.NR - 2];
#pragma unroll
		for (int k = 0; k < K8; ++k) {
			T acc = 0.0;
#pragma unroll
			for (int qq = 0; qq < NF; ++qq) acc += Pg[(f.rev ? NF - 1 - qq : qq) * f.NR + j] * Q[qq * K8 + k];
			acc *= o.winv[f.NR - 2][j];
			YR[k * NF + j] = acc;
			a.z[f.zR + kNV * f.NR + k * f.NR + j] = acc;
		}
	}
	wave_sync();
	T hL[K8], hR[K8];
	if (ownL) side_face_gradient<NF>(f.kL, f.NL, j, tL, kL, scL, YL, hL);
	if (ownR) side_face_gradient<NF>(f.kR, f.NR, j, tR, kR, scR, YR, hR);
	wave_sync();   // every y is read
	if (ownL)
#pragma unroll
		for (int k = 0; k < K8; ++k) YL[k * NF + j] = hL[k];
	if (ownR)
#pragma unroll
		for (int k = 0; k < K8; ++k) YR[k * NF + j] = hR[k];
	wave_sync();
	T fq[kNV];
	if (point) {   // the flux at this face point: inviscid first, then each side's face gradient P h and viscous flux
		const T Jf = sqrt(nj0 * nj0 + nj1 * nj1);
		const T nu[2] = { nj0 / Jf, nj1 / Jf };
		T nnf[kNV], fvL[K8], fvR[K8];
		if constexpr (FLUX == DGX_FLUX_LF) numflux_lf<T, 2>(sL, sR, nu, nnf);
		else numflux_roe<T, 2>(sL, sR, nu, nnf);
		{
			T g[K8];
			const T* P = o.Pg[f.NL - 2];
#pragma unroll
			for (int k = 0; k < K8; ++k) {
				T acc = 0.0;
#pragma unroll
				for (int i = 0; i < NF; ++i)
					if (i < f.NL) acc += P[q * f.NL + i] * YL[k * NF + i];
				g[k] = acc;
			}
			ns_viscous_flux2<T>(sL, g, a.mu, a.Pr, fvL);
			// boundary faces: the boundary gradient is the interior one (grad_from_internal)
			if (!inter) ns_viscous_flux2<T>(sR, g, a.mu, a.Pr, fvR);
		}
		if (inter) {
			T g[K8];
			const T* P = o.Pg[f.NR - 2];
#pragma unroll
			for (int k = 0; k < K8; ++k) {
				T acc = 0.0;
#pragma unroll
				for (int i = 0; i < NF; ++i)
					if (i < f.NR) acc += P[qr * f.NR + i] * YR[k * NF + i];
				g[k] = acc;
			}
			ns_viscous_flux2<T>(sR, g, a.mu, a.Pr, fvR);
		}
		const T wj = o.w[q] * Jf;
		// no-slip flux walls: the energy flux is the provided one (numerical_flux_central_T.c:222-230)
		const bool nf_provided = f.bc == DGX_BC_NOSLIP_ADIABATIC || f.bc == DGX_BC_NOSLIP_DIABATIC;
#pragma unroll
		for (int v = 0; v < kNV; ++v) {
			const T vis = nu[0] * T(0.5) * (fvL[0 + 2 * v] + fvR[0 + 2 * v]) + nu[1] * T(0.5) * (fvL[1 + 2 * v] + fvR[1 + 2 * v]);
			fq[v] = wj * ((nf_provided && v == kNV - 1) ? T(a.walls.nfE) : nnf[v] + vis);
		}
	}
	// every q was read before the y step's sync
	if (point)
#pragma unroll
		for (int v = 0; v < kNV; ++v) Q[q * kNV + v] = fq[v];
	wave_sync();
	if (ownL) lift_side<NF>(o, f.NL, j, false, false, Q, a.z + f.zL);
	if (ownR) lift_side<NF>(o, f.NR, j, f.rev != 0, true, Q, a.z + f.zR);
}

// a TRI group's buffers in the precision T (FP32: the shadows of the geometry and mass data; the per-evaluation
// traces and gradients use the first half of the FP64-sized buffers)
template <typename T>
TriArgsT<T> tri_args(const TriGroup& g, const TriOps<T>* ops, const T* u, T* r, const T* z)
{
	TriArgsT<T> a{};
	a.u = u + g.u_off;
	a.r = r ? r + g.u_off : nullptr;
	a.z = z ? z + g.z_off : nullptr;
	a.tr = reinterpret_cast<T*>(g.tr);
	a.gv = reinterpret_cast<T*>(g.gv);
	a.gvt = reinterpret_cast<T*>(g.gvt);
	if constexpr (std::is_same<T, double>::value) {
		a.met = g.met;
		a.minv = g.minv;
		a.jinv = g.jinv;
	} else {
		a.met = g.met32;
		a.minv = g.minv32;
		a.jinv = g.jinv32;
	}
	a.ops = ops + g.p;
	a.E = g.E;
	a.affine = g.affine;
	return a;
}

// one pass of batches over the group (the kernels loop over batches when the grid is capped)
int tri_grid(int E, int epb)
{
	return std::max(1, std::min((E + epb - 1) / epb, 16384));
}

template <typename T>
MixArgsT<T> mix_args(const HpData& H, const Ctx& c, const T* u)
{
	constexpr bool f64 = std::is_same<T, double>::value;
	MixArgsT<T> a{};
	a.faces = H.mfaces;
	a.u = u;
	if constexpr (f64) {
		a.nJ = H.mnJ;
		a.z = H.z64;
		a.hops = H.ops64;
		a.tops = H.tri_ops64;
	} else {
		a.nJ = H.mnJ32;
		a.z = H.z32;
		a.hops = H.ops32;
		a.tops = H.tri_ops32;
	}
	for (size_t t = 0; t < H.tri.size() && t < (size_t)kTriMaxP; ++t) {
		const TriGroup& g = H.tri[t];
		a.ttr[t] = reinterpret_cast<const T*>(g.tr);
		a.tgvt[t] = reinterpret_cast<const T*>(g.gvt);
		if constexpr (f64) {
			a.tK[t] = g.affine ? nullptr : g.K;
			a.tjinv[t] = g.jinv;
		} else {
			a.tK[t] = g.affine ? nullptr : g.K32;
			a.tjinv[t] = g.jinv32;
		}
	}
	for (size_t g = 0; g < H.groups.size(); ++g) {   // FP32 uses the first half of the FP64-sized buffers
		const Ctx& s = H.groups[g]->c;
		a.qgvt[s.N] = reinterpret_cast<const T*>(s.ns_gvt);
		a.qkap[s.N] = reinterpret_cast<const T*>(s.ns_kappa);
	}
	for (int k = 0; k < 5; ++k) a.bc_prim[k] = c.cfg.bc_prim[k];
	a.mu = c.cfg.mu;
	a.Pr = c.cfg.Pr;
	a.walls = NsWalls{ c.cfg.wall_nf_E, c.cfg.wall_omega, c.cfg.wall_rho, c.cfg.wall_E };
	a.wxy = H.mxy;
	return a;
}

// the faces with a TRI side, one launch per run of equal Nf, spread over the mesh's side streams
template <typename T>
hipError_t launch_mix(const HpData& H, const Ctx& c, const T* u, bool ns)
{
	MixArgsT<T> a = mix_args<T>(H, c, u);
	const bool lf = c.cfg.num_flux == DGX_FLUX_LF;
	for (size_t i = 0; i < H.mruns.size(); ++i) {
		const MixRun& run = H.mruns[i];
		a.start = run.start;
		a.count = run.count;
		const hipStream_t st = H.lane[H.mrun_lane[i]];
#define DGX_MIX(NFV)                                                                                           \
	case NFV: {                                                                                                \
		const dim3 grid((run.count + mix_fpb(NFV) - 1) / mix_fpb(NFV)), block(64 * kMixWaves);                 \
		if (ns) {                                                                                              \
			if (lf) hipLaunchKernelGGL((k_mixf_ns<T, NFV, DGX_FLUX_LF>), grid, block, 0, st, a);               \
			else hipLaunchKernelGGL((k_mixf_ns<T, NFV, DGX_FLUX_ROE>), grid, block, 0, st, a);                 \
		} else {                                                                                               \
			if (lf) hipLaunchKernelGGL((k_mixf<T, NFV, DGX_FLUX_LF>), grid, block, 0, st, a);                  \
			else hipLaunchKernelGGL((k_mixf<T, NFV, DGX_FLUX_ROE>), grid, block, 0, st, a);                    \
		}                                                                                                      \
		break;                                                                                                 \
	}
		switch (run.Nf) {
			DGX_MIX(2) DGX_MIX(3) DGX_MIX(4) DGX_MIX(5) DGX_MIX(6)
		default: return hipErrorInvalidValue;
		}
#undef DGX_MIX
		const hipError_t e = hipGetLastError();
		if (e != hipSuccess) return e;
	}
	return hipSuccess;
}

}  // namespace

#define DGX_TRI_SWITCH(KERNEL, EPB_OF)                                                                     \
	switch (g.p) {                                                                                         \
	case 1: hipLaunchKernelGGL(KERNEL<1>, dim3(tri_grid(g.E, EPB_OF(1))), dim3(256), 0, st, a); break;     \
	case 2: hipLaunchKernelGGL(KERNEL<2>, dim3(tri_grid(g.E, EPB_OF(2))), dim3(256), 0, st, a); break;     \
	case 3: hipLaunchKernelGGL(KERNEL<3>, dim3(tri_grid(g.E, EPB_OF(3))), dim3(256), 0, st, a); break;     \
	case 4: hipLaunchKernelGGL(KERNEL<4>, dim3(tri_grid(g.E, EPB_OF(4))), dim3(256), 0, st, a); break;     \
	case 5: hipLaunchKernelGGL(KERNEL<5>, dim3(tri_grid(g.E, EPB_OF(5))), dim3(256), 0, st, a); break;     \
	default: return hipErrorInvalidValue;                                                                  \
	}
// the same for the kernels templated on the scalar type
#define DGX_TRI_SWITCH_T(KERNEL, T, EPB_OF)                                                                    \
	switch (g.p) {                                                                                             \
	case 1: hipLaunchKernelGGL((KERNEL<T, 1>), dim3(tri_grid(g.E, EPB_OF(1))), dim3(256), 0, st, a); break;    \
	case 2: hipLaunchKernelGGL((KERNEL<T, 2>), dim3(tri_grid(g.E, EPB_OF(2))), dim3(256), 0, st, a); break;    \
	case 3: hipLaunchKernelGGL((KERNEL<T, 3>), dim3(tri_grid(g.E, EPB_OF(3))), dim3(256), 0, st, a); break;    \
	case 4: hipLaunchKernelGGL((KERNEL<T, 4>), dim3(tri_grid(g.E, EPB_OF(4))), dim3(256), 0, st, a); break;    \
	case 5: hipLaunchKernelGGL((KERNEL<T, 5>), dim3(tri_grid(g.E, EPB_OF(5))), dim3(256), 0, st, a); break;    \
	default: return hipErrorInvalidValue;                                                                      \
	}

hipError_t launch_tri_trace64(const TriGroup& g, const TriOps<double>* ops, const double* u, hipStream_t st)
{
	if (g.E == 0) return hipSuccess;
	TriArgs a = tri_args<double>(g, ops, u, nullptr, nullptr);
	DGX_TRI_SWITCH_T(k_tri_trace, double, epb_trace)
	return hipGetLastError();
}

hipError_t launch_tri_trace32(const TriGroup& g, const TriOps<float>* ops, const float* u, hipStream_t st)
{
	if (g.E == 0) return hipSuccess;
	TriArgsT<float> a = tri_args<float>(g, ops, u, nullptr, nullptr);
	DGX_TRI_SWITCH_T(k_tri_trace, float, epb_trace)
	return hipGetLastError();
}

// TRI residual kernel: the matrix-core form unless DGX_TRI_KERNEL=valu (the VALU form, for comparison)
static bool tri_use_mfma()
{
	static const bool mfma = [] {
		const char* e = std::getenv("DGX_TRI_KERNEL");
		return !(e && std::strcmp(e, "valu") == 0);
	}();
	return mfma;
}

constexpr int epb_mfma(int) { return 16; }   // four wavefronts of four elements

hipError_t launch_tri_rhs64(const TriGroup& g, const TriOps<double>* ops, const double* u, double* r,
                            const double* z, hipStream_t st)
{
	if (g.E == 0) return hipSuccess;
	TriArgs a = tri_args<double>(g, ops, u, r, z);
	if (tri_use_mfma()) {
		DGX_TRI_SWITCH(k_tri_rhs_mfma, epb_mfma)
	} else {
		DGX_TRI_SWITCH_T(k_tri_rhs, double, epb_rhs)
	}
	return hipGetLastError();
}

hipError_t launch_tri_rhs32(const TriGroup& g, const TriOps<float>* ops, const float* u, float* r, const float* z,
                            hipStream_t st)
{
	if (g.E == 0) return hipSuccess;
	TriArgsT<float> a = tri_args<float>(g, ops, u, r, z);
	DGX_TRI_SWITCH_T(k_tri_rhs, float, epb_rhs)
	return hipGetLastError();
}

hipError_t launch_tri_ns_phase64(const TriGroup& g, const TriOps<double>* ops, const double* u, double* r,
                                 const double* z, double mu, double Pr, int phase, hipStream_t st)
{
	if (g.E == 0) return hipSuccess;
	TriArgs a = tri_args<double>(g, ops, u, r, z);
	a.mu = mu;
	a.Pr = Pr;
	if (phase == 1) {
		if (g.affine && tri_use_mfma()) {
			DGX_TRI_SWITCH(k_tri_gv_mfma, epb_mfma)
		} else {
			DGX_TRI_SWITCH_T(k_tri_gv, double, epb_gv)
		}
	} else if (g.affine && tri_use_mfma()) {
		DGX_TRI_SWITCH(k_tri_ns_mfma, epb_mfma)
	} else {
		DGX_TRI_SWITCH_T(k_tri_ns, double, epb_ns)
	}
	return hipGetLastError();
}

hipError_t launch_tri_ns_phase32(const TriGroup& g, const TriOps<float>* ops, const float* u, float* r,
                                 const float* z, double mu, double Pr, int phase, hipStream_t st)
{
	if (g.E == 0) return hipSuccess;
	TriArgsT<float> a = tri_args<float>(g, ops, u, r, z);
	a.mu = mu;
	a.Pr = Pr;
	if (phase == 1) {
		DGX_TRI_SWITCH_T(k_tri_gv, float, epb_gv)
	} else {
		DGX_TRI_SWITCH_T(k_tri_ns, float, epb_ns)
	}
	return hipGetLastError();
}

hipError_t launch_tri_minv64(const TriGroup& g, const TriOps<double>* ops, double* r, unsigned long long* dmax,
                             hipStream_t st)
{
	if (g.E == 0) return hipSuccess;
	TriArgs a = tri_args<double>(g, ops, r, r, nullptr);
	switch (g.p) {
	case 1: hipLaunchKernelGGL(k_tri_minv<1>, dim3(tri_grid(g.E, epb_minv(1))), dim3(256), 0, st, a, dmax); break;
	case 2: hipLaunchKernelGGL(k_tri_minv<2>, dim3(tri_grid(g.E, epb_minv(2))), dim3(256), 0, st, a, dmax); break;
	case 3: hipLaunchKernelGGL(k_tri_minv<3>, dim3(tri_grid(g.E, epb_minv(3))), dim3(256), 0, st, a, dmax); break;
	case 4: hipLaunchKernelGGL(k_tri_minv<4>, dim3(tri_grid(g.E, epb_minv(4))), dim3(256), 0, st, a, dmax); break;
	case 5: hipLaunchKernelGGL(k_tri_minv<5>, dim3(tri_grid(g.E, epb_minv(5))), dim3(256), 0, st, a, dmax); break;
	default: return hipErrorInvalidValue;
	}
	return hipGetLastError();
}

#undef DGX_TRI_SWITCH
#undef DGX_TRI_SWITCH_T

hipError_t launch_mix_faces64(const HpData& H, const Ctx& c, const double* u, hipStream_t)
{
	return launch_mix(H, c, u, false);
}

hipError_t launch_mix_faces_ns64(const HpData& H, const Ctx& c, const double* u, hipStream_t)
{
	return launch_mix(H, c, u, true);
}

hipError_t launch_mix_faces32(const HpData& H, const Ctx& c, const float* u, bool ns)
{
	return launch_mix(H, c, u, ns);
}

}  // namespace dgx
