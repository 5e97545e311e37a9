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

template <typename S>
hipError_t run_hp_faces(Ctx& c, const S* u, S* z, const S* nJ, const HpOps<S>* ops)
{
	const HpData& H = *c.hp;
	if (H.Fnc == 0) return hipSuccess;
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
