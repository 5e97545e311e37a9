// k_hp.hip -- faces between elements of different orders on 2-D variable-order meshes.
//
// A face takes the larger order of its two volumes (face_solver.c:123-131), so its cubature is the GL rule of
// Nf = max(NL, NR) points.  One thread per face:
//   1. gathers both sides' GLL face nodes (NL and NR per variable) from the state,
//   2. interpolates them to the Nf face GL points (get_operator__cv0_vs_fc_T with the face order,
//      compute_face_rlhs_T.c:67-80), the right side read in the left ordering,
//   3. evaluates the numerical flux (Roe-Pike / Lax-Friedrichs) with the stored n J_f and scales by w_q J_f,
//   4. lifts it to each side's own GL face points, z(a) = sum_q P[q][a] g(q) / w_a with P the exact
//      interpolation from the element's N GL points to the Nf face points: the residual kernel's GL-space face
//      lift Lift(c) w_a z(a) / w_a then equals tw0_vt_fc (get_operator__tw0_vt_fc_T,
//      compute_face_rlhs_T.c:51-64) of the element's own order; the right side is negated in its own ordering
//      (finalize_face_rhs_dg_like_T, compute_face_rlhs_T.c:347-359).
// The residual kernel of each order reads z for its faces marked kBcNonconforming instead of evaluating them.
#include "dgx_internal.h"
#include "dgx_device.h"

namespace dgx {
namespace {

template <typename S>
struct HpFaceArgs {
	const S* u;
	S* z;
	const HpFace* faces;
	const S* nJ;
	const HpOps<S>* ops;
	int F;
};

__device__ inline int face_node_2d(int N, int lf, int i)
{
	const int fix = (lf & 1) ? N - 1 : 0;
	return (lf >> 1) == 0 ? fix + N * i : i + N * fix;
}

template <typename S, int FLUX>
__global__ void __launch_bounds__(64) k_hp_faces(HpFaceArgs<S> a)
{
	constexpr int NV = 4, NM = kHpMaxN;
	const int f = blockIdx.x * blockDim.x + threadIdx.x;
	if (f >= a.F) return;
	const HpFace hf = a.faces[f];
	const int NL = hf.info & 7, NR = (hf.info >> 3) & 7, lfL = (hf.info >> 6) & 3, lfR = (hf.info >> 8) & 3;
	const bool flip = (hf.info >> 11) & 1;   // ord bit 1: the right face runs the other way
	const int Nf = NL > NR ? NL : NR;
	const HpOps<S>* o = a.ops;
	S uL[NV][NM], uR[NV][NM];
#pragma unroll
	for (int i = 0; i < NM; ++i) {
		if (i < NL) {
			const S* p = a.u + hf.uL + face_node_2d(NL, lfL, i);
#pragma unroll
			for (int v = 0; v < NV; ++v) uL[v][i] = p[v * NL * NL];
		}
		if (i < NR) {
			const S* p = a.u + hf.uR + face_node_2d(NR, lfR, i);
#pragma unroll
			for (int v = 0; v < NV; ++v) uR[v][i] = p[v * NR * NR];
		}
	}
	S g[NM][NV];
#pragma unroll
	for (int q = 0; q < NM; ++q) {
		if (q < Nf) {
			const int qr = flip ? Nf - 1 - q : q;
			const S* IL = o->Ig[NL][Nf] + q * NL;
			const S* IR = o->Ig[NR][Nf] + qr * NR;
			S sL[NV], sR[NV];
#pragma unroll
			for (int v = 0; v < NV; ++v) {
				S aL = S(0), aR = S(0);
#pragma unroll
				for (int i = 0; i < NM; ++i) {
					if (i < NL) aL += IL[i] * uL[v][i];
					if (i < NR) aR += IR[i] * uR[v][i];
				}
				sL[v] = aL;
				sR[v] = aR;
			}
			const S* nj = a.nJ + ((size_t)f * NM + q) * 2;
			const S Jf = tsqrt(nj[0] * nj[0] + nj[1] * nj[1]);
			const S Jfi = trcp(Jf);
			const S nu[2] = { nj[0] * Jfi, nj[1] * Jfi };
			S nnf[NV];
			if constexpr (FLUX == DGX_FLUX_LF) numflux_lf<S, 2>(sL, sR, nu, nnf);
			else numflux_roe<S, 2>(sL, sR, nu, nnf);
			const S wj = o->w[Nf][q] * Jf;
#pragma unroll
			for (int v = 0; v < NV; ++v) g[q][v] = wj * nnf[v];
		}
	}
	// lifts to each side's own GL face points (the right side in its own ordering, negated)
#pragma unroll
	for (int p = 0; p < NM; ++p) {
		if (p < NL) {
#pragma unroll
			for (int v = 0; v < NV; ++v) {
				S acc = S(0);
#pragma unroll
				for (int q = 0; q < NM; ++q)
					if (q < Nf) acc += o->Pg[NL][Nf][q * NL + p] * g[q][v];
				a.z[hf.zL + v * NL + p] = acc * o->winv[NL][p];
			}
		}
		if (p < NR) {
#pragma unroll
			for (int v = 0; v < NV; ++v) {
				S acc = S(0);
#pragma unroll
				for (int q = 0; q < NM; ++q)
					if (q < Nf) acc += o->Pg[NR][Nf][(flip ? Nf - 1 - q : q) * NR + p] * g[q][v];
				a.z[hf.zR + v * NR + p] = S(0) - acc * o->winv[NR][p];
			}
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
	S* z;
	const HpFace* faces;
	const S* nJ;
	const HpOps<S>* ops;
	const S* gvt[kHpMaxN + 1];   // each order's face traces of G_v (E_g x 4 x 2 x 4 x N) and kappa (E_g x 4 x N)
	const S* kap[kHpMaxN + 1];
	S mu, Pr;
	int F;
};

template <typename S, int FLUX, int NL, int NR>
__global__ void __launch_bounds__(64) k_hp_faces_ns(HpFaceNsArgs<S> a, int f0, int nf)
{
	constexpr int NV = 4, FPB = 8, Nf = NL > NR ? NL : NR;
	constexpr S eta = S(1.01 * 4.0);
	// this pair's operators, staged once per block
	__shared__ S IgL[Nf * NL], IgR[Nf * NR], PgL[Nf * NL], PgR[Nf * NR], wf[Nf], wiL[NL], wiR[NR];
	// per face, aliased across the stages: the GLL face nodes (stage 1-2), then h = trace G_v + eta kappa y at
	// the own face nodes (stage 3-4); w q (stage 2-3), then w J_f F-hat (stage 4-5)
	constexpr int HS = 2 * NV * (NL + NR);
	constexpr int QS = 2 * NV * Nf;
	__shared__ S Hs[FPB][HS];
	__shared__ S Qs[FPB][QS];
	const int slot = threadIdx.x >> 3, j = threadIdx.x & 7;
	for (int i = threadIdx.x; i < Nf * NL; i += 64) {
		IgL[i] = a.ops->Ig[NL][Nf][i];
		PgL[i] = a.ops->Pg[NL][Nf][i];
	}
	for (int i = threadIdx.x; i < Nf * NR; i += 64) {
		IgR[i] = a.ops->Ig[NR][Nf][i];
		PgR[i] = a.ops->Pg[NR][Nf][i];
	}
	if (threadIdx.x < Nf) wf[threadIdx.x] = a.ops->w[Nf][threadIdx.x];
	if (threadIdx.x < NL) wiL[threadIdx.x] = a.ops->winv[NL][threadIdx.x];
	if (threadIdx.x < NR) wiR[threadIdx.x] = a.ops->winv[NR][threadIdx.x];
	const int fl = blockIdx.x * FPB + slot;
	const bool valid = fl < nf;
	const int f = f0 + (valid ? fl : 0);
	const HpFace hf = a.faces[f];
	const int lfL = (hf.info >> 6) & 3, lfR = (hf.info >> 8) & 3;
	const bool flip = (hf.info >> 11) & 1;   // ord bit 1: the right face runs the other way
	S* uLs = Hs[slot];                       // [v][i], i < NL
	S* uRs = Hs[slot] + NV * NL;             // [v][i], i < NR
	if (valid && j < NL) {
#pragma unroll
		for (int v = 0; v < NV; ++v) uLs[v * NL + j] = a.u[hf.uL + v * NL * NL + face_node_2d(NL, lfL, j)];
	}
	if (valid && j < NR) {
#pragma unroll
		for (int v = 0; v < NV; ++v) uRs[v * NR + j] = a.u[hf.uR + v * NR * NR + face_node_2d(NR, lfR, j)];
	}
	__syncthreads();
	// traces at this lane's face point (left ordering), kept in registers; w q into LDS
	S sL[NV], sR[NV];
	S* qf = Qs[slot];                        // [q][d * NV + v]
	if (valid && j < Nf) {
		const int q = j, qr = flip ? Nf - 1 - q : q;
		const S* nj = a.nJ + ((size_t)f * kHpMaxN + q) * 2;
		const S wq = wf[q];
#pragma unroll
		for (int v = 0; v < NV; ++v) {
			S aL = S(0), aR = S(0);
#pragma unroll
			for (int i = 0; i < NL; ++i) aL += IgL[q * NL + i] * uLs[v * NL + i];
#pragma unroll
			for (int i = 0; i < NR; ++i) aR += IgR[qr * NR + i] * uRs[v * NR + i];
			sL[v] = aL;
			sR[v] = aR;
			const S dl = S(0.5) * (aR - aL);
			qf[q * 2 * NV + v] = wq * (nj[0] * dl);
			qf[q * 2 * NV + NV + v] = wq * (nj[1] * dl);
		}
	}
	__syncthreads();
	// projections y to each side's own GL face points (written out for the residual kernels), and
	// h = trace G_v + eta kappa y at the lane's own node (the face nodes in LDS are dead)
	S hL[2 * NV], hR[2 * NV];
	if (valid && j < NL) {
		const S* t = a.gvt[NL] + ((size_t)hf.eL * 4 + lfL) * 2 * NV * NL;
		const S kap = a.kap[NL][((size_t)hf.eL * 4 + lfL) * NL + j];
#pragma unroll
		for (int k = 0; k < 2 * NV; ++k) {
			S acc = S(0);
#pragma unroll
			for (int q = 0; q < Nf; ++q) acc += PgL[q * NL + j] * qf[q * 2 * NV + k];
			acc = acc * wiL[j];
			a.z[hf.zL + NV * NL + k * NL + j] = acc;
			hL[k] = t[k * NL + j] + eta * kap * acc;
		}
	}
	if (valid && j < NR) {
		const S* t = a.gvt[NR] + ((size_t)hf.eR * 4 + lfR) * 2 * NV * NR;
		const S kap = a.kap[NR][((size_t)hf.eR * 4 + lfR) * NR + j];
#pragma unroll
		for (int k = 0; k < 2 * NV; ++k) {
			S acc = S(0);
#pragma unroll
			for (int q = 0; q < Nf; ++q) acc += PgR[(flip ? Nf - 1 - q : q) * NR + j] * qf[q * 2 * NV + k];
			acc = acc * wiR[j];
			a.z[hf.zR + NV * NR + k * NR + j] = acc;
			hR[k] = t[k * NR + j] + eta * kap * acc;
		}
	}
	__syncthreads();   // w q and the face nodes are dead
	S* hLs = Hs[slot];                       // [k][i], i < NL
	S* hRs = Hs[slot] + 2 * NV * NL;         // [k][i], i < NR
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
	S* gq = Qs[slot];                        // [q][v]
	if (valid && j < Nf) {
		const int q = j, qr = flip ? Nf - 1 - q : q;
		S gL[2 * NV], gR[2 * NV];
#pragma unroll
		for (int k = 0; k < 2 * NV; ++k) {
			S aL = S(0), aR = S(0);
#pragma unroll
			for (int i = 0; i < NL; ++i) aL += PgL[q * NL + i] * hLs[k * NL + i];
#pragma unroll
			for (int i = 0; i < NR; ++i) aR += PgR[qr * NR + i] * hRs[k * NR + i];
			gL[k] = aL;
			gR[k] = aR;
		}
		const S* nj = a.nJ + ((size_t)f * kHpMaxN + q) * 2;
		const S Jf = tsqrt(nj[0] * nj[0] + nj[1] * nj[1]);
		const S Jfi = trcp(Jf);
		const S nu[2] = { nj[0] * Jfi, nj[1] * Jfi };
		S nnf[NV], fvL[2 * NV], fvR[2 * NV];
		if constexpr (FLUX == DGX_FLUX_LF) numflux_lf<S, 2>(sL, sR, nu, nnf);
		else numflux_roe<S, 2>(sL, sR, nu, nnf);
		ns_viscous_flux2<S>(sL, gL, a.mu, a.Pr, fvL);
		ns_viscous_flux2<S>(sR, gR, a.mu, a.Pr, fvR);
		const S wj = wf[q] * Jf;
#pragma unroll
		for (int v = 0; v < NV; ++v) {
			const S vis = nu[0] * S(0.5) * (fvL[0 + 2 * v] + fvR[0 + 2 * v]) + nu[1] * S(0.5) * (fvL[1 + 2 * v] + fvR[1 + 2 * v]);
			gq[q * NV + v] = wj * (nnf[v] + vis);
		}
	}
	__syncthreads();
	// lifts to each side's own GL face points (the right side negated)
	if (valid && j < NL) {
#pragma unroll
		for (int v = 0; v < NV; ++v) {
			S acc = S(0);
#pragma unroll
			for (int q = 0; q < Nf; ++q) acc += PgL[q * NL + j] * gq[q * NV + v];
			a.z[hf.zL + v * NL + j] = acc * wiL[j];
		}
	}
	if (valid && j < NR) {
#pragma unroll
		for (int v = 0; v < NV; ++v) {
			S acc = S(0);
#pragma unroll
			for (int q = 0; q < Nf; ++q) acc += PgR[(flip ? Nf - 1 - q : q) * NR + j] * gq[q * NV + v];
			a.z[hf.zR + v * NR + j] = S(0) - acc * wiR[j];
		}
	}
}

template <typename S>
hipError_t run_hp_faces(Ctx& c, const S* u, S* z, const S* nJ, const HpOps<S>* ops)
{
	const HpData& H = *c.hp;
	if (H.Fnc == 0) return hipSuccess;
	if (c.cfg.pde == DGX_PDE_NAVIER_STOKES) {
		HpFaceNsArgs<S> a{};
		a.u = u;
		a.z = z;
		a.faces = H.faces;
		a.nJ = nJ;
		a.ops = ops;
		for (size_t g = 0; g < H.groups.size(); ++g) {
			const Ctx& s = H.groups[g]->c;
			a.gvt[s.N] = reinterpret_cast<const S*>(s.ns_gvt);   // FP32 uses the first half of the FP64 buffers
			a.kap[s.N] = reinterpret_cast<const S*>(s.ns_kappa);
		}
		a.mu = S(c.cfg.mu);
		a.Pr = S(c.cfg.Pr);
		a.F = H.Fnc;
		const bool lf = c.cfg.num_flux == DGX_FLUX_LF;
		for (const HpPair& pr : H.pairs) {
			const int blocks = (pr.count + 7) / 8;
#define DGX_HP_PAIR(L, R)                                                                                              \
	case L * 8 + R:                                                                                                    \
		if (lf) hipLaunchKernelGGL((k_hp_faces_ns<S, DGX_FLUX_LF, L, R>), dim3(blocks), dim3(64), 0, c.stream, a, pr.start, pr.count); \
		else hipLaunchKernelGGL((k_hp_faces_ns<S, DGX_FLUX_ROE, L, R>), dim3(blocks), dim3(64), 0, c.stream, a, pr.start, pr.count); \
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
	HpFaceArgs<S> a{ u, z, H.faces, nJ, ops, H.Fnc };
	const int blocks = (H.Fnc + 63) / 64;
	if (c.cfg.num_flux == DGX_FLUX_LF) hipLaunchKernelGGL((k_hp_faces<S, DGX_FLUX_LF>), dim3(blocks), dim3(64), 0, c.stream, a);
	else hipLaunchKernelGGL((k_hp_faces<S, DGX_FLUX_ROE>), dim3(blocks), dim3(64), 0, c.stream, a);
	return hipGetLastError();
}

}  // namespace

hipError_t launch_hp_faces64(Ctx& c, const double* u)
{
	return run_hp_faces<double>(c, u, c.hp->z64, c.hp->nJ, c.hp->ops64);
}

hipError_t launch_hp_faces32(Ctx& c, const float* u)
{
	return run_hp_faces<float>(c, u, c.hp->z32, c.hp->nJ32, c.hp->ops32);
}

}  // namespace dgx
