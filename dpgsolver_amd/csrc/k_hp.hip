// k_hp.hip -- faces between elements of different orders on 2-D variable-order meshes.
//
// A face takes the larger order of its two volumes (face_solver.c:123-131), so its cubature is the GL rule of
// Nf = max(NL, NR) points.  Per face (one kernel instance per (NL, NR) pair, the faces sorted by pair at upload,
// eight lanes per face):
//   1. both sides' GLL face nodes (NL and NR per variable) are gathered from the state,
//   2. interpolated to the Nf face GL points (get_operator__cv0_vs_fc_T with the face order,
//      compute_face_rlhs_T.c:67-80), the right side read in the left ordering,
//   3. the numerical flux (Roe-Pike / Lax-Friedrichs) is evaluated with the stored n J_f and scaled by w_q J_f,
//   4. and lifted to each side's own GL face points, z(a) = sum_q P[q][a] g(q) / w_a with P the exact
//      interpolation from the element's N GL points to the Nf face points: the residual kernel's GL-space face
//      lift Lift(c) w_a z(a) / w_a then equals tw0_vt_fc (get_operator__tw0_vt_fc_T,
//      compute_face_rlhs_T.c:51-64) of the element's own order; the right side is negated in its own ordering
//      (finalize_face_rhs_dg_like_T, compute_face_rlhs_T.c:347-359).
// The residual kernel of each order reads z for its faces marked kBcNonconforming instead of evaluating them.
//
// Mixed precision (PERT): the state is read in FP64 and every trace, flux and lift is formed in FP64; each
// side's lift subtracts the normal Euler flux of its element's reference state c_e (the element's first GLL
// node) on that element's own face geometry (n J)_own, as the mixed residual kernels do for their own faces,
// so the constant part cancels exactly against their volume term; z (and the BR2 products y) are then rounded
// to FP32 for the FP32 group kernels.
#include <type_traits>

#include "dgx_internal.h"
#include "dgx_device.h"

namespace dgx {
namespace {

// S: the storage type of z (and y); the FP64 pointers serve the mixed form
template <typename S>
struct HpFaceArgs {
	const S* u;
	const double* u64;
	S* z;
	const HpFace* faces;
	const S* nJ;
	const double* nJ64;
	const double* nJhi;   // mixed form: n J of the higher-order side's own face
	const HpOps<S>* ops;
	const HpOps<double>* ops64;
	int F;
};

__device__ inline int face_node_2d(int N, int lf, int i)
{
	const int fix = (lf & 1) ? N - 1 : 0;
	return (lf >> 1) == 0 ? fix + N * i : i + N * fix;
}

// n . F(c) for the reference state c and a (not necessarily unit) normal n: with n = (n J)_own of an element's own
// face geometry, the face integral of this constant part cancels the element's volume integral of F(c) exactly
// (its free-stream preservation), which the geometry of a face between orders does not give on curved meshes
template <typename C>
__device__ inline void ref_normal_flux(const C* c, const C* n, C* fn)
{
	const C ci = trcp(c[0]);
	const C cu[2] = { c[1] * ci, c[2] * ci };
	const C pc = C(kGM1) * (c[3] - C(0.5) * c[0] * (cu[0] * cu[0] + cu[1] * cu[1]));
	euler_normal_flux<C, 2>(c, cu, pc, n, fn);
}

// Euler: one instance per (NL, NR) pair, eight lanes per face: the lanes load the face nodes, evaluate the flux
// at one face point each and lift to one own face node each, through LDS.
template <typename S, int FLUX, int NL, int NR, bool PERT>
__global__ void __launch_bounds__(64) k_hp_faces(HpFaceArgs<S> a, int f0, int nf)
{
	using C = typename std::conditional<PERT, double, S>::type;
	constexpr int NV = 4, FPB = 8, Nf = NL > NR ? NL : NR;
	const C* U;
	const C* NJ;
	const HpOps<C>* o;
	if constexpr (PERT) { U = a.u64; NJ = a.nJ64; o = a.ops64; }
	else { U = a.u; NJ = a.nJ; o = a.ops; }
	__shared__ C IgL[Nf * NL], IgR[Nf * NR], PgL[Nf * NL], PgR[Nf * NR], wf[Nf], wiL[NL], wiR[NR];
	__shared__ C uLs[FPB][NV * NL], uRs[FPB][NV * NR];   // GLL face nodes [v][i], own orderings
	__shared__ C gqL[FPB][Nf * NV], gqR[FPB][Nf * NV];   // w J_f F-hat per side (the mixed form: minus F(c_e))
	const int slot = threadIdx.x >> 3, j = threadIdx.x & 7;
	for (int i = threadIdx.x; i < Nf * NL; i += 64) {
		IgL[i] = o->Ig[NL][Nf][i];
		PgL[i] = o->Pg[NL][Nf][i];
	}
	for (int i = threadIdx.x; i < Nf * NR; i += 64) {
		IgR[i] = o->Ig[NR][Nf][i];
		PgR[i] = o->Pg[NR][Nf][i];
	}
	if (threadIdx.x < Nf) wf[threadIdx.x] = o->w[Nf][threadIdx.x];
	if (threadIdx.x < NL) wiL[threadIdx.x] = o->winv[NL][threadIdx.x];
	if (threadIdx.x < NR) wiR[threadIdx.x] = o->winv[NR][threadIdx.x];
	const int fl = blockIdx.x * FPB + slot;
	const bool valid = fl < nf;
	const int f = f0 + (valid ? fl : 0);
	const HpFace hf = a.faces[f];
	const int lfL = (hf.info >> 6) & 3, lfR = (hf.info >> 8) & 3;
	const bool flip = (hf.info >> 11) & 1;   // ord bit 1: the right face runs the other way
	if (valid && j < NL) {
#pragma unroll
		for (int v = 0; v < NV; ++v) uLs[slot][v * NL + j] = U[hf.uL + v * NL * NL + face_node_2d(NL, lfL, j)];
	}
	if (valid && j < NR) {
#pragma unroll
		for (int v = 0; v < NV; ++v) uRs[slot][v * NR + j] = U[hf.uR + v * NR * NR + face_node_2d(NR, lfR, j)];
	}
	__syncthreads();
	if (valid && j < Nf) {
		const int q = j, qr = flip ? Nf - 1 - q : q;
		C sL[NV], sR[NV];
#pragma unroll
		for (int v = 0; v < NV; ++v) {
			C aL = C(0), aR = C(0);
#pragma unroll
			for (int i = 0; i < NL; ++i) aL += IgL[q * NL + i] * uLs[slot][v * NL + i];
#pragma unroll
			for (int i = 0; i < NR; ++i) aR += IgR[qr * NR + i] * uRs[slot][v * NR + i];
			sL[v] = aL;
			sR[v] = aR;
		}
		const C* nj = NJ + ((size_t)f * kHpMaxN + q) * 2;
		const C Jf = tsqrt(nj[0] * nj[0] + nj[1] * nj[1]);
		const C Jfi = trcp(Jf);
		const C nu[2] = { nj[0] * Jfi, nj[1] * Jfi };
		C nnf[NV];
		if constexpr (FLUX == DGX_FLUX_LF) numflux_lf<C, 2>(sL, sR, nu, nnf);
		else numflux_roe<C, 2>(sL, sR, nu, nnf);
		const C wj = wf[q] * Jf;
		if constexpr (PERT) {
			C cL[NV], cR[NV], fcL[NV], fcR[NV];
#pragma unroll
			for (int v = 0; v < NV; ++v) {
				cL[v] = U[hf.uL + v * NL * NL];
				cR[v] = U[hf.uR + v * NR * NR];
			}
			const double* nh = a.nJhi + ((size_t)f * kHpMaxN + q) * 2;
			ref_normal_flux<C>(cL, (NL < NR) ? nj : nh, fcL);   // each side's own n J (left orientation)
			ref_normal_flux<C>(cR, (NL < NR) ? nh : nj, fcR);
#pragma unroll
			for (int v = 0; v < NV; ++v) {
				gqL[slot][q * NV + v] = wj * nnf[v] - wf[q] * fcL[v];
				gqR[slot][q * NV + v] = wj * nnf[v] - wf[q] * fcR[v];
			}
		} else {
#pragma unroll
			for (int v = 0; v < NV; ++v) gqL[slot][q * NV + v] = gqR[slot][q * NV + v] = wj * nnf[v];
		}
	}
	__syncthreads();
	// lifts to each side's own GL face points (the right side negated, in its own ordering)
	if (valid && j < NL) {
#pragma unroll
		for (int v = 0; v < NV; ++v) {
			C acc = C(0);
#pragma unroll
			for (int q = 0; q < Nf; ++q) acc += PgL[q * NL + j] * gqL[slot][q * NV + v];
			a.z[hf.zL + v * NL + j] = S(acc * wiL[j]);
		}
	}
	if (valid && j < NR) {
#pragma unroll
		for (int v = 0; v < NV; ++v) {
			C acc = C(0);
#pragma unroll
			for (int q = 0; q < Nf; ++q) acc += PgR[(flip ? Nf - 1 - q : q) * NR + j] * gqR[slot][q * NV + v];
			a.z[hf.zR + v * NR + j] = S(C(0) - acc * wiR[j]);
		}
	}
}

// ------------------------------------------------------------------------------------------- Navier-Stokes
// BR2 at a face between orders (compute_grad_coef_dg_T.c:171-307, compute_face_rlhs_dg_T.c:243-300): both sides
// use q = n J_f (s_R - s_L) / 2 at the Nf face points; side e's face lift of the gradient is, at its own GL
// nodes, Lift(c) y_e(a) / (w_c J(c, a)) with the projection y_e(a) = sum_q P_e[q][a] w_q q(q) / w_a, and its
// trace at the face is kappa_e(a) y_e(a).  The face gradient at the Nf points is therefore
// P_e (trace G_v + eta kappa_e y_e), and the face flux is the inviscid flux plus the central
// n . (fv_L + fv_R) / 2 (numerical_flux_central_T.c:76-104).  Per element face the kernel writes the lifted
// flux z [v][a] and then y [d][v][a], which the residual kernel of each order reads in place of its own face
// traces.  One instance per (NL, NR) pair (the faces arrive sorted by pair), eight lanes per face (lane = face
// point or own face node); the pair's operators, the face nodes and the cross-lane stages go through LDS.
template <typename S>
struct HpFaceNsArgs {
	const S* u;
	const double* u64;
	S* z;
	const HpFace* faces;
	const S* nJ;
	const double* nJ64;
	const double* nJhi;          // mixed form: n J of the higher-order side's own face
	const HpOps<S>* ops;
	const HpOps<double>* ops64;
	const S* gvt[kHpMaxN + 1];   // each order's face traces of G_v (E_g x 4 x 2 x 4 x N) and kappa (E_g x 4 x N)
	const S* kap[kHpMaxN + 1];
	double mu, Pr;
	int F;
};

template <typename S, int FLUX, int NL, int NR, bool PERT>
__global__ void __launch_bounds__(64) k_hp_faces_ns(HpFaceNsArgs<S> a, int f0, int nf)
{
	using C = typename std::conditional<PERT, double, S>::type;
	constexpr int NV = 4, FPB = 8, Nf = NL > NR ? NL : NR;
	constexpr C eta = C(1.01 * 4.0);
	const C* U;
	const C* NJ;
	const HpOps<C>* o;
	if constexpr (PERT) { U = a.u64; NJ = a.nJ64; o = a.ops64; }
	else { U = a.u; NJ = a.nJ; o = a.ops; }
	const C mu = C(a.mu), Pr = C(a.Pr);
	// this pair's operators, staged once per block
	__shared__ C IgL[Nf * NL], IgR[Nf * NR], PgL[Nf * NL], PgR[Nf * NR], wf[Nf], wiL[NL], wiR[NR];
	// per face, aliased across the stages: the GLL face nodes (stage 1-2), then h = trace G_v + eta kappa y at
	// the own face nodes (stage 3-4); w q (stage 2-3), then w J_f F-hat for each side (stage 4-5)
	constexpr int HS = 2 * NV * (NL + NR);
	constexpr int QS = 2 * NV * Nf;
	__shared__ C Hs[FPB][HS];
	__shared__ C Qs[FPB][QS];
	const int slot = threadIdx.x >> 3, j = threadIdx.x & 7;
	for (int i = threadIdx.x; i < Nf * NL; i += 64) {
		IgL[i] = o->Ig[NL][Nf][i];
		PgL[i] = o->Pg[NL][Nf][i];
	}
	for (int i = threadIdx.x; i < Nf * NR; i += 64) {
		IgR[i] = o->Ig[NR][Nf][i];
		PgR[i] = o->Pg[NR][Nf][i];
	}
	if (threadIdx.x < Nf) wf[threadIdx.x] = o->w[Nf][threadIdx.x];
	if (threadIdx.x < NL) wiL[threadIdx.x] = o->winv[NL][threadIdx.x];
	if (threadIdx.x < NR) wiR[threadIdx.x] = o->winv[NR][threadIdx.x];
	const int fl = blockIdx.x * FPB + slot;
	const bool valid = fl < nf;
	const int f = f0 + (valid ? fl : 0);
	const HpFace hf = a.faces[f];
	const int lfL = (hf.info >> 6) & 3, lfR = (hf.info >> 8) & 3;
	const bool flip = (hf.info >> 11) & 1;   // ord bit 1: the right face runs the other way
	C* uLs = Hs[slot];                       // [v][i], i < NL
	C* uRs = Hs[slot] + NV * NL;             // [v][i], i < NR
	if (valid && j < NL) {
#pragma unroll
		for (int v = 0; v < NV; ++v) uLs[v * NL + j] = U[hf.uL + v * NL * NL + face_node_2d(NL, lfL, j)];
	}
	if (valid && j < NR) {
#pragma unroll
		for (int v = 0; v < NV; ++v) uRs[v * NR + j] = U[hf.uR + v * NR * NR + face_node_2d(NR, lfR, j)];
	}
	__syncthreads();
	// traces at this lane's face point (left ordering), kept in registers; w q into LDS
	C sL[NV], sR[NV];
	C* qf = Qs[slot];                        // [q][d * NV + v]
	if (valid && j < Nf) {
		const int q = j, qr = flip ? Nf - 1 - q : q;
		const C* nj = NJ + ((size_t)f * kHpMaxN + q) * 2;
		const C wq = wf[q];
#pragma unroll
		for (int v = 0; v < NV; ++v) {
			C aL = C(0), aR = C(0);
#pragma unroll
			for (int i = 0; i < NL; ++i) aL += IgL[q * NL + i] * uLs[v * NL + i];
#pragma unroll
			for (int i = 0; i < NR; ++i) aR += IgR[qr * NR + i] * uRs[v * NR + i];
			sL[v] = aL;
			sR[v] = aR;
			const C dl = C(0.5) * (aR - aL);
			qf[q * 2 * NV + v] = wq * (nj[0] * dl);
			qf[q * 2 * NV + NV + v] = wq * (nj[1] * dl);
		}
	}
	__syncthreads();
	// projections y to each side's own GL face points (written out for the residual kernels), and
	// h = trace G_v + eta kappa y at the lane's own node (the face nodes in LDS are dead)
	C hL[2 * NV], hR[2 * NV];
	if (valid && j < NL) {
		const S* t = a.gvt[NL] + ((size_t)hf.eL * 4 + lfL) * 2 * NV * NL;
		const C kap = C(a.kap[NL][((size_t)hf.eL * 4 + lfL) * NL + j]);
#pragma unroll
		for (int k = 0; k < 2 * NV; ++k) {
			C acc = C(0);
#pragma unroll
			for (int q = 0; q < Nf; ++q) acc += PgL[q * NL + j] * qf[q * 2 * NV + k];
			acc = acc * wiL[j];
			a.z[hf.zL + NV * NL + k * NL + j] = S(acc);
			hL[k] = C(t[k * NL + j]) + eta * kap * acc;
		}
	}
	if (valid && j < NR) {
		const S* t = a.gvt[NR] + ((size_t)hf.eR * 4 + lfR) * 2 * NV * NR;
		const C kap = C(a.kap[NR][((size_t)hf.eR * 4 + lfR) * NR + j]);
#pragma unroll
		for (int k = 0; k < 2 * NV; ++k) {
			C acc = C(0);
#pragma unroll
			for (int q = 0; q < Nf; ++q) acc += PgR[(flip ? Nf - 1 - q : q) * NR + j] * qf[q * 2 * NV + k];
			acc = acc * wiR[j];
			a.z[hf.zR + NV * NR + k * NR + j] = S(acc);
			hR[k] = C(t[k * NR + j]) + eta * kap * acc;
		}
	}
	__syncthreads();   // w q and the face nodes are dead
	C* hLs = Hs[slot];                       // [k][i], i < NL
	C* hRs = Hs[slot] + 2 * NV * NL;         // [k][i], i < NR
	if (valid && j < NL) {
#pragma unroll
		for (int k = 0; k < 2 * NV; ++k) hLs[k * NL + j] = hL[k];
	}
	if (valid && j < NR) {
#pragma unroll
		for (int k = 0; k < 2 * NV; ++k) hRs[k * NR + j] = hR[k];
	}
	__syncthreads();
	// face gradients P_e h_e and the face flux at this lane's point
	C* gqL = Qs[slot];                       // [q][v]
	C* gqR = Qs[slot] + NV * Nf;             // [q][v] (differs from gqL in the mixed form only)
	if (valid && j < Nf) {
		const int q = j, qr = flip ? Nf - 1 - q : q;
		C gL[2 * NV], gR[2 * NV];
#pragma unroll
		for (int k = 0; k < 2 * NV; ++k) {
			C aL = C(0), aR = C(0);
#pragma unroll
			for (int i = 0; i < NL; ++i) aL += PgL[q * NL + i] * hLs[k * NL + i];
#pragma unroll
			for (int i = 0; i < NR; ++i) aR += PgR[qr * NR + i] * hRs[k * NR + i];
			gL[k] = aL;
			gR[k] = aR;
		}
		const C* nj = NJ + ((size_t)f * kHpMaxN + q) * 2;
		const C Jf = tsqrt(nj[0] * nj[0] + nj[1] * nj[1]);
		const C Jfi = trcp(Jf);
		const C nu[2] = { nj[0] * Jfi, nj[1] * Jfi };
		C nnf[NV], fvL[2 * NV], fvR[2 * NV];
		if constexpr (FLUX == DGX_FLUX_LF) numflux_lf<C, 2>(sL, sR, nu, nnf);
		else numflux_roe<C, 2>(sL, sR, nu, nnf);
		ns_viscous_flux2<C>(sL, gL, mu, Pr, fvL);
		ns_viscous_flux2<C>(sR, gR, mu, Pr, fvR);
		const C wj = wf[q] * Jf;
		C fcL[NV], fcR[NV];
		if constexpr (PERT) {
			C cL[NV], cR[NV];
#pragma unroll
			for (int v = 0; v < NV; ++v) {
				cL[v] = U[hf.uL + v * NL * NL];
				cR[v] = U[hf.uR + v * NR * NR];
			}
			const double* nh = a.nJhi + ((size_t)f * kHpMaxN + q) * 2;
			ref_normal_flux<C>(cL, (NL < NR) ? nj : nh, fcL);   // each side's own n J (left orientation)
			ref_normal_flux<C>(cR, (NL < NR) ? nh : nj, fcR);
		}
#pragma unroll
		for (int v = 0; v < NV; ++v) {
			const C vis = nu[0] * C(0.5) * (fvL[0 + 2 * v] + fvR[0 + 2 * v]) + nu[1] * C(0.5) * (fvL[1 + 2 * v] + fvR[1 + 2 * v]);
			const C tot = nnf[v] + vis;
			if constexpr (PERT) {
				gqL[q * NV + v] = wj * tot - wf[q] * fcL[v];
				gqR[q * NV + v] = wj * tot - wf[q] * fcR[v];
			} else {
				gqL[q * NV + v] = gqR[q * NV + v] = wj * tot;
			}
		}
	}
	__syncthreads();
	// lifts to each side's own GL face points (the right side negated)
	if (valid && j < NL) {
#pragma unroll
		for (int v = 0; v < NV; ++v) {
			C acc = C(0);
#pragma unroll
			for (int q = 0; q < Nf; ++q) acc += PgL[q * NL + j] * gqL[q * NV + v];
			a.z[hf.zL + v * NL + j] = S(acc * wiL[j]);
		}
	}
	if (valid && j < NR) {
#pragma unroll
		for (int v = 0; v < NV; ++v) {
			C acc = C(0);
#pragma unroll
			for (int q = 0; q < Nf; ++q) acc += PgR[(flip ? Nf - 1 - q : q) * NR + j] * gqR[q * NV + v];
			a.z[hf.zR + v * NR + j] = S(C(0) - acc * wiR[j]);
		}
	}
}

template <typename S, bool PERT>
hipError_t run_hp_faces(Ctx& c, const S* u, const double* u64, S* z, const S* nJ, const HpOps<S>* ops)
{
	const HpData& H = *c.hp;
	if (H.Fnc == 0) return hipSuccess;
	if (c.cfg.pde == DGX_PDE_NAVIER_STOKES) {
		HpFaceNsArgs<S> a{};
		a.u = u;
		a.u64 = u64;
		a.z = z;
		a.faces = H.faces;
		a.nJ = nJ;
		a.nJ64 = H.nJ;
		a.nJhi = H.nJhi;
		a.ops = ops;
		a.ops64 = H.ops64;
		for (size_t g = 0; g < H.groups.size(); ++g) {
			const Ctx& s = H.groups[g]->c;
			a.gvt[s.N] = reinterpret_cast<const S*>(s.ns_gvt);   // FP32 uses the first half of the FP64 buffers
			a.kap[s.N] = reinterpret_cast<const S*>(s.ns_kappa);
		}
		a.mu = c.cfg.mu;
		a.Pr = c.cfg.Pr;
		a.F = H.Fnc;
		const bool lf = c.cfg.num_flux == DGX_FLUX_LF;
		for (size_t ip = 0; ip < H.pairs.size(); ++ip) {
			const HpPair& pr = H.pairs[ip];
			const hipStream_t st = H.lane[H.pair_lane[ip]];
			const int blocks = (pr.count + 7) / 8;
#define DGX_HP_PAIR(L, R)                                                                                              \
	case L * 8 + R:                                                                                                    \
		if (lf) hipLaunchKernelGGL((k_hp_faces_ns<S, DGX_FLUX_LF, L, R, PERT>), dim3(blocks), dim3(64), 0, st, a, pr.start, pr.count); \
		else hipLaunchKernelGGL((k_hp_faces_ns<S, DGX_FLUX_ROE, L, R, PERT>), dim3(blocks), dim3(64), 0, st, a, pr.start, pr.count); \
		break;
			switch (pr.NL * 8 + pr.NR) {
				DGX_HP_PAIR(2, 3) DGX_HP_PAIR(2, 4) DGX_HP_PAIR(2, 5) DGX_HP_PAIR(2, 6)
				DGX_HP_PAIR(3, 2) DGX_HP_PAIR(3, 4) DGX_HP_PAIR(3, 5) DGX_HP_PAIR(3, 6)
				DGX_HP_PAIR(4, 2) DGX_HP_PAIR(4, 3) DGX_HP_PAIR(4, 5) DGX_HP_PAIR(4, 6)
				DGX_HP_PAIR(5, 2) DGX_HP_PAIR(5, 3) DGX_HP_PAIR(5, 4) DGX_HP_PAIR(5, 6)
				DGX_HP_PAIR(6, 2) DGX_HP_PAIR(6, 3) DGX_HP_PAIR(6, 4) DGX_HP_PAIR(6, 5)
			default: return hipErrorInvalidValue;
			}
#undef DGX_HP_PAIR
			DGX_HIP_CHECK(hipGetLastError());
		}
		return hipSuccess;
	}
	HpFaceArgs<S> a{ u, u64, z, H.faces, nJ, H.nJ, H.nJhi, ops, H.ops64, H.Fnc };
	const bool lf = c.cfg.num_flux == DGX_FLUX_LF;
	for (size_t ip = 0; ip < H.pairs.size(); ++ip) {
		const HpPair& pr = H.pairs[ip];
		const hipStream_t st = H.lane[H.pair_lane[ip]];
		const int blocks = (pr.count + 7) / 8;
#define DGX_HP_PAIR(L, R)                                                                                              \
	case L * 8 + R:                                                                                                    \
		if (lf) hipLaunchKernelGGL((k_hp_faces<S, DGX_FLUX_LF, L, R, PERT>), dim3(blocks), dim3(64), 0, st, a, pr.start, pr.count); \
		else hipLaunchKernelGGL((k_hp_faces<S, DGX_FLUX_ROE, L, R, PERT>), dim3(blocks), dim3(64), 0, st, a, pr.start, pr.count); \
		break;
		switch (pr.NL * 8 + pr.NR) {
			DGX_HP_PAIR(2, 3) DGX_HP_PAIR(2, 4) DGX_HP_PAIR(2, 5) DGX_HP_PAIR(2, 6)
			DGX_HP_PAIR(3, 2) DGX_HP_PAIR(3, 4) DGX_HP_PAIR(3, 5) DGX_HP_PAIR(3, 6)
			DGX_HP_PAIR(4, 2) DGX_HP_PAIR(4, 3) DGX_HP_PAIR(4, 5) DGX_HP_PAIR(4, 6)
			DGX_HP_PAIR(5, 2) DGX_HP_PAIR(5, 3) DGX_HP_PAIR(5, 4) DGX_HP_PAIR(5, 6)
			DGX_HP_PAIR(6, 2) DGX_HP_PAIR(6, 3) DGX_HP_PAIR(6, 4) DGX_HP_PAIR(6, 5)
		default: return hipErrorInvalidValue;
		}
#undef DGX_HP_PAIR
		DGX_HIP_CHECK(hipGetLastError());
	}
	return hipSuccess;
}

}  // namespace

// Fork and join of the side streams around one phase.  Both also run inside graph captures (dgx_graph_create_rhs),
// so each event is recorded on this call, in the current capture, before any stream waits on it: a wait on an
// event whose last record belonged to an earlier capture (whose graph may since have been destroyed) would make
// the capture depend on nodes of that other graph.
hipError_t hp_fork(HpData& H, hipStream_t main)
{
	DGX_HIP_CHECK(hipEventRecord(H.fork_ev, main));
	for (int k = 0; k < HpData::kLanes; ++k) DGX_HIP_CHECK(hipStreamWaitEvent(H.lane[k], H.fork_ev, 0));
	return hipSuccess;
}

hipError_t hp_join(HpData& H, hipStream_t main)
{
	for (int k = 0; k < HpData::kLanes; ++k) {
		DGX_HIP_CHECK(hipEventRecord(H.join_ev[k], H.lane[k]));
		DGX_HIP_CHECK(hipStreamWaitEvent(main, H.join_ev[k], 0));
	}
	return hipSuccess;
}

// The face kernels between orders, one launch per (NL, NR) pair on that pair's side stream; the caller forks
// the side streams before and joins them after.
hipError_t launch_hp_faces64(Ctx& c, const double* u)
{
	return run_hp_faces<double, false>(c, u, nullptr, c.hp->z64, c.hp->nJ, c.hp->ops64);
}

hipError_t launch_hp_faces32(Ctx& c, const float* u)
{
	return run_hp_faces<float, false>(c, u, nullptr, c.hp->z32, c.hp->nJ32, c.hp->ops32);
}

hipError_t launch_hp_faces_mixed(Ctx& c, const double* u)
{
	return run_hp_faces<float, true>(c, nullptr, u, c.hp->z32, c.hp->nJ32, c.hp->ops32);
}

}  // namespace dgx
