// k_ns.hip -- 2-D Navier-Stokes residual with BR2 viscous terms (QUAD; interior, Riemann and no-slip wall faces),
// FP64 and FP32, and its linearisation in dual numbers (J v, diagonal-block columns)
// (FP32: the FP32 state, geometry shadow and operator tables, as the FP32 Euler residual), and the mixed
// form (PERT: FP64 state in, FP32 contractions on u - c_e, FP64 pointwise fluxes, as the Euler mixed kernel).
//
// The weak gradient (compute_grad_coef_dg_T.c:171-307, 'e' branch) is formed at the volume cubature (GL)
// nodes.  With M = Ibar^T W J Ibar, M^-1 Ibar^T W x = Ibar^-1 (x / J), so the projected volume gradient is the
// pointwise G_v = (J grad s) / J, and the face term M^-1 tw0_vt_fc (J_f n (s_num - s)) is the residual kernel's
// face lift divided by W J:  G_f(c, a) = Lift_side(c) q(a) / (w_c J(c, a)),  q = J_f n (s_nb - s_own) / 2
// (central s_num).  The volume flux uses G = G_v + sum_faces G_f at the GL nodes (cv0_vs_vc grad_coef =
// Ibar Ibar^-1 G); each side of a face uses the trace of G_v + 1.01 NFMAX G_f(face) (NFMAX = 4 in 2-D;
// constructor_partial_grad_fc_interp, compute_face_rlhs_dg_T.c:243-300).  The trace of GL nodal values at
// xi = -1 / +1 is Lift_side = I^-T e_side, so the face part of that trace is q(a) kappa(a) with the geometric
// factor kappa(a) = sum_c Lift_side(c)^2 / (w_c J(c, a)).
//
//   k_ns_gv:  per element, the face traces of G_v and kappa (its own state and geometry only).
//   k_rhs_ns: G_v again, the face products q from the own and neighbour traces, G for the volume flux
//     F_euler(s) + fv(s, G) (fv the negated viscous flux, compute_Flux_navier_stokes_0,
//     flux_navier_stokes_T.c:1005-1040), the weak divergence, and per face the inviscid numerical flux plus the
//     central n . (fv_L + fv_R) / 2 (numerical_flux_central_T.c:76-104) of the two sides' face gradients
//     trace(G_v) + eta kappa q, scaled by J_f and lifted.  Both sides of a face evaluate it from identical
//     inputs in the left ordering (the traces of G_v and kappa are read back from k_ns_gv's output by both).
//     A boundary face takes the boundary state (ns_wall_state) as its right trace, so q = J_f n (s_b - s_L) / 2
//     (compute_g_coef_related_boundary, compute_grad_coef_dg_T.c:527-575), the interior face gradient as the
//     boundary one, and on the no-slip flux walls the provided energy flux nf_E in place of the energy
//     equation's numerical flux.
#include <type_traits>

#include "dgx_internal.h"
#include "dgx_device.h"

namespace dgx {

namespace {

constexpr int kNsNV = 4;

// elements per workgroup of up to 128 lanes (dual numbers at p <= 2: up to 64 lanes, for the LDS)
template <typename S, int N>
constexpr int ns_epb()
{
	constexpr int lanes = (sizeof(S) > 8 && N <= 3) ? 64 : 128;
	return (lanes / (N * N)) > 0 ? lanes / (N * N) : 1;
}

// what the kernels compute: the residual, J v (dual numbers seeded with v; with mass_scale, J v - s_e M_e v),
// one column of every element's diagonal block dR_e/du_e (dual numbers seeded with the local dof col, the
// neighbours' traces and gradient traces held fixed), or one column of every face block dR_e/du_n, n the
// neighbour across local face blockIdx.y (the neighbour's dof col seeded, through its trace and its gradient
// trace on that face; the element's own state and the other neighbours held fixed)
constexpr int kNsRhs = 0, kNsJvp = 1, kNsDiag = 2, kNsOffd = 3;

// node of local face lf at face position a (faces r-, r+, s-, s+)
template <int N>
__device__ inline int ns_face_node(int lf, int a)
{
	const int fix = (lf & 1) ? N - 1 : 0;
	return (lf >> 1) == 0 ? fix + N * a : a + N * fix;
}

template <typename S>
struct NsArgs {
	using G = typename RealOf<S>::type;   // precision of the state, geometry and operators (dual numbers: FP64)
	const G* u;
	const double* u64;     // PERT: FP64 state (u unused)
	const G* v;            // kNsJvp: the direction
	G* out;           // residual; kNsJvp: J v; kNsDiag: the diagonal blocks, E x nb x nb
	S* gvt;           // E x 4 x 2 x 4 x N: traces of G_v on the faces (own face ordering)
	S* kap;           // E x 4 x N: kappa per face node
	const G* met;     // E x 4 x Np
	const G* detJ;    // E x Np
	const G* nJ;      // F x N x 2
	const ElemFace* ef;
	const Ops1D<G>* ops;
	S mu, Pr;
	int sutherland;             // viscosity by Sutherland's formula: selects the k_rhs_ns<..., SUTH = true> kernels
	double r_s;
	int E;
	int col;                    // kNsDiag, kNsOffd: the seeded local dof
	int offd_transposed;        // kNsOffd: out[((e 4 + lf) nb + col) nb + row] (else [.. row) nb + col])
	const double* mass_scale;   // kNsJvp: s_e of the -s_e M_e v term (or NULL)
	// variable-order meshes: per element face to an element of another order, the lifted face flux [v][a]
	// and then the projected face products q [d][v][a] (own GL face points and ordering)
	const S* zf;
	// boundary faces: the Riemann free stream, the no-slip wall data and (rotating walls) the face GL point
	// coordinates, F x N x 2
	double bc_prim[5];
	NsWalls walls;
	const double* wall_xy;
};


// shared per-element set-up: operators, the GLL state, its GL values and (k_rhs_ns) the GL face traces
template <typename S, int N>
struct NsBlock {
	static constexpr int Np = N * N, EPB = ns_epb<S, N>();
	S I[N * N], DhT[N * N], Lift[2][N], w[N];
	int perm[8][N];
	static constexpr int UR = (kNsNV * Np > 4 * kNsNV * N) ? kNsNV * Np : 4 * kNsNV * N;
	S U[EPB][UR];              // GLL nodal state (k_rhs_ns: then the face fluxes)
	S V[EPB][kNsNV * Np];      // GL values
	double C[EPB][kNsNV];      // PERT: the element's reference state c_e (its first GLL node)
	ElemFace ef[EPB][4];
};

// PERT: U = u - c_e from the FP64 state (the constant part has no gradient and its fluxes cancel exactly by
// free-stream preservation), so FP32 rounding acts only on the within-element variation
template <typename S, int N, bool PERT, int MODE>
__device__ inline void ns_load(NsBlock<S, N>& B, const NsArgs<S>& a, int sl, int t, int e, bool valid)
{
	constexpr int Np = N * N;
	for (int i = threadIdx.x; i < N * N; i += blockDim.x) {
		B.I[i] = a.ops->I[i];
		B.DhT[i] = a.ops->DhT[i];
	}
	for (int i = threadIdx.x; i < N; i += blockDim.x) {
		B.Lift[0][i] = a.ops->Lift[0][i];
		B.Lift[1][i] = a.ops->Lift[1][i];
		B.w[i] = a.ops->w[i];
	}
	for (int i = threadIdx.x; i < 8 * N; i += blockDim.x) B.perm[i / N][i % N] = a.ops->perm[i / N][i % N];
	if (valid) {
		if (t < 4) B.ef[sl][t] = a.ef[(size_t)e * 4 + t];
		if constexpr (PERT) {
			if (t < kNsNV) B.C[sl][t] = a.u64[((size_t)e * kNsNV + t) * Np];
			for (int v = 0; v < kNsNV; ++v)
				B.U[sl][v * Np + t] = S(a.u64[((size_t)e * kNsNV + v) * Np + t] - a.u64[((size_t)e * kNsNV + v) * Np]);
		} else {
			for (int v = 0; v < kNsNV; ++v) {
				const size_t gi = ((size_t)e * kNsNV + v) * Np + t;
				if constexpr (MODE == kNsJvp) B.U[sl][v * Np + t] = S(a.u[gi], a.v[gi]);
				else if constexpr (MODE == kNsDiag) B.U[sl][v * Np + t] = S(a.u[gi], (v * Np + t == a.col) ? 1.0 : 0.0);
				else B.U[sl][v * Np + t] = a.u[gi];
			}
		}
	}
	__syncthreads();
	// GL values: I along r, then along s (through V as scratch)
	S tmp[kNsNV];
	const int c0 = t % N, c1 = t / N;
	if (valid) {
#pragma unroll
		for (int v = 0; v < kNsNV; ++v) {
			S acc = S(0.0);
#pragma unroll
			for (int m = 0; m < N; ++m) acc += B.I[c0 * N + m] * B.U[sl][v * Np + m + N * c1];
			tmp[v] = acc;
		}
#pragma unroll
		for (int v = 0; v < kNsNV; ++v) B.V[sl][v * Np + t] = tmp[v];
	}
	__syncthreads();
	if (valid) {
#pragma unroll
		for (int v = 0; v < kNsNV; ++v) {
			S acc = S(0.0);
#pragma unroll
			for (int m = 0; m < N; ++m) acc += B.I[c1 * N + m] * B.V[sl][v * Np + c0 + N * m];
			tmp[v] = acc;
		}
	}
	__syncthreads();
	if (valid)
#pragma unroll
		for (int v = 0; v < kNsNV; ++v) B.V[sl][v * Np + t] = tmp[v];
	__syncthreads();
}

// G_v = (J grad s) / J at GL node t (c0, c1): g[v + 4 d]; m[(i + 2 j)] = J dxi_j / dx_i
template <typename S, int N>
__device__ inline void ns_volume_gradient(const NsBlock<S, N>& B, const NsArgs<S>& a, int sl, int t, int e, S* g)
{
	constexpr int Np = N * N;
	const int c0 = t % N, c1 = t / N;
	const typename NsArgs<S>::G* m = a.met + (size_t)e * 4 * Np;
	const S m00 = m[0 * Np + t], m10 = m[1 * Np + t], m01 = m[2 * Np + t], m11 = m[3 * Np + t];
	const S Ji = trcp(a.detJ[(size_t)e * Np + t]);
#pragma unroll
	for (int v = 0; v < kNsNV; ++v) {
		S d0 = S(0.0), d1 = S(0.0);   // d/dxi_0, d/dxi_1 (GL-collocated derivative Dh[i][j] = DhT[j N + i])
#pragma unroll
		for (int mm = 0; mm < N; ++mm) {
			d0 += B.DhT[mm * N + c0] * B.V[sl][v * Np + mm + N * c1];
			d1 += B.DhT[mm * N + c1] * B.V[sl][v * Np + c0 + N * mm];
		}
		g[v + 4 * 0] = (m00 * d0 + m01 * d1) * Ji;
		g[v + 4 * 1] = (m10 * d0 + m11 * d1) * Ji;
	}
}

template <typename S, int N, bool PERT, int MODE>
__global__ void __launch_bounds__(128) k_ns_gv(NsArgs<S> a)
{
	constexpr int Np = N * N, EPB = ns_epb<S, N>();
	__shared__ NsBlock<S, N> B;
	__shared__ S Gv[EPB][2 * kNsNV * Np];
	__shared__ S WJi[EPB][Np];   // 1 / (w_c0 w_c1 J)
	const int sl = threadIdx.x / Np, t = threadIdx.x % Np;
	const int e = blockIdx.x * EPB + sl;
	const bool valid = sl < EPB && e < a.E;
	const int slc = sl < EPB ? sl : 0;
	ns_load<S, N, PERT, MODE>(B, a, slc, t, valid ? e : 0, valid);
	if (valid) {
		S g[2 * kNsNV];
		ns_volume_gradient<S, N>(B, a, slc, t, e, g);
#pragma unroll
		for (int k = 0; k < 2 * kNsNV; ++k) Gv[slc][(k / kNsNV * kNsNV + k % kNsNV) * Np + t] = g[k];
		WJi[slc][t] = trcp(B.w[t % N] * B.w[t / N] * a.detJ[(size_t)e * Np + t]);
	}
	__syncthreads();
	if (valid) {
		// traces of G_v: items (lf, d, v, a) in the output order
		for (int item = t; item < 4 * 2 * kNsNV * N; item += Np) {
			const int at = item % N, r = item / N, v = r % kNsNV, d = (r / kNsNV) % 2, lf = r / (2 * kNsNV);
			const int fd = lf >> 1, es = lf & 1;   // normal direction and end (xi = -1 / +1) of the face
			S acc = S(0.0);
#pragma unroll
			for (int c = 0; c < N; ++c) acc += B.Lift[es][c] * Gv[slc][(d * kNsNV + v) * Np + (fd == 0 ? c + N * at : at + N * c)];
			a.gvt[(size_t)e * 4 * 2 * kNsNV * N + item] = acc;
		}
		// kappa(a) = sum_c Lift(c)^2 / (w_c J(c, a)) = sum_c Lift(c)^2 w_a WJi(c, a)
		for (int item = t; item < 4 * N; item += Np) {
			const int at = item % N, lf = item / N, fd = lf >> 1, es = lf & 1;
			S acc = S(0.0);
#pragma unroll
			for (int c = 0; c < N; ++c)
				acc += B.Lift[es][c] * B.Lift[es][c] * B.w[at] * WJi[slc][fd == 0 ? c + N * at : at + N * c];
			a.kap[(size_t)e * 4 * N + item] = acc;
		}
	}
}

template <typename S, int N, int FLUX, bool PERT, int MODE, bool SUTH>
__global__ void __launch_bounds__(128) k_rhs_ns(NsArgs<S> a)
{
	constexpr int Np = N * N, EPB = ns_epb<S, N>();
	const S eta = S(1.01 * 4.0);
	__shared__ NsBlock<S, N> B;
	// LDS shared between phases (occupancy): the face traces, dead once the face fluxes are formed, hold W fr_d
	// for the volume divergence; the GLL state, dead once the traces are formed, holds the face fluxes
	constexpr int TXS = (8 * kNsNV * N > 2 * kNsNV * Np) ? 8 * kNsNV * N : 2 * kNsNV * Np;
	__shared__ S TX[EPB][TXS];
	__shared__ S QF[EPB][4 * 2 * kNsNV * N];  // q = J_f n_d (s_nb - s_own) / 2 per face, own ordering
	S (*TRo)[TXS] = TX;                        // own face traces at the face GL nodes, own ordering: [0, 16 N)
	// neighbour face traces, the neighbour's ordering: TX[slot] + 16 N
	S (*X)[TXS] = TX;                          // W fr_d at the GL nodes (after the face phase)
	constexpr int UR = NsBlock<S, N>::UR;
	S (*RES)[UR] = B.U;                        // J_f nnf per face, own ordering (after the traces)
	const int sl = threadIdx.x / Np, t = threadIdx.x % Np;
	const int e = blockIdx.x * EPB + sl;
	const bool valid = sl < EPB && e < a.E;
	const int slc = sl < EPB ? sl : 0;
	ns_load<S, N, PERT, MODE>(B, a, slc, t, valid ? e : 0, valid);
	const int c0 = t % N, c1 = t / N;
	if (valid) {
		// face traces: GLL face values (own from LDS, neighbour gathered), then I along the face
		for (int item = t; item < 4 * N; item += Np) {
			const int lf = item / N, i = item % N;
			const ElemFace f = B.ef[slc][lf];
			const int fn = f.packed & 15, bc = (f.packed >> 12) & 15;
			const int en = f.en >= 0 ? f.en : e;   // boundary faces: the own element (its value is replaced below)
			S* nbt = TX[slc] + 4 * kNsNV * N;
			S aos[kNsNV];
#pragma unroll
			for (int v = 0; v < kNsNV; ++v) {
				S ao = S(0.0), an = S(0.0);
#pragma unroll
				for (int b = 0; b < N; ++b) {
					ao += B.I[i * N + b] * B.U[slc][v * Np + ns_face_node<N>(lf, b)];
					const size_t gn = ((size_t)en * kNsNV + v) * Np + ns_face_node<N>(fn, b);
					if constexpr (PERT) an += B.I[i * N + b] * S(a.u64[gn] - B.C[slc][v]);
					else if constexpr (MODE == kNsJvp) an += B.I[i * N + b] * S(a.u[gn], a.v[gn]);
					else if constexpr (MODE == kNsOffd)   // the seeded face's neighbour dof col
						an += B.I[i * N + b] *
						      S(a.u[gn], (lf == (int)blockIdx.y && v * Np + ns_face_node<N>(fn, b) == a.col) ? 1.0 : 0.0);
					else an += B.I[i * N + b] * S(a.u[gn]);   // kNsDiag: the neighbour held fixed
				}
				TRo[slc][(lf * kNsNV + v) * N + i] = ao;
				nbt[(lf * kNsNV + v) * N + i] = an;
				aos[v] = ao;
			}
			if (f.en < 0 && bc != kBcNonconforming) {
				// boundary face: the boundary state from the own trace (PERT: in FP64 from c_e + delta, back to a
				// delta; dual numbers: with its derivative)
				const typename NsArgs<S>::G* nJ = a.nJ + ((size_t)f.f * N + i) * 2;
				const double Jd = sqrt((double)nJ[0] * nJ[0] + (double)nJ[1] * nJ[1]);
				double xy[2] = { 0.0, 0.0 };
				if (bc == DGX_BC_NOSLIP_ROTATING) {
					xy[0] = a.wall_xy[((size_t)f.f * N + i) * 2];
					xy[1] = a.wall_xy[((size_t)f.f * N + i) * 2 + 1];
				}
				if constexpr (PERT) {
					double sl[kNsNV], sb[kNsNV];
					const double nd[2] = { (double)nJ[0] / Jd, (double)nJ[1] / Jd };
#pragma unroll
					for (int v = 0; v < kNsNV; ++v) sl[v] = B.C[slc][v] + (double)aos[v];
					ns_wall_state<double>(bc, sl, nd, a.bc_prim, a.walls, xy, sb);
#pragma unroll
					for (int v = 0; v < kNsNV; ++v) nbt[(lf * kNsNV + v) * N + i] = S(sb[v] - B.C[slc][v]);
				} else {
					S sb[kNsNV];
					const S ns2[2] = { S((double)nJ[0] / Jd), S((double)nJ[1] / Jd) };
					ns_wall_state<S>(bc, aos, ns2, a.bc_prim, a.walls, xy, sb);
#pragma unroll
					for (int v = 0; v < kNsNV; ++v) nbt[(lf * kNsNV + v) * N + i] = sb[v];
				}
			}
		}
	}
	__syncthreads();
	if (valid) {
		// face products in the left ordering, stored at the own index
		for (int item = t; item < 4 * N; item += Np) {
			const int lf = item / N, i = item % N;
			const ElemFace f = B.ef[slc][lf];
			const int side = (f.packed >> 4) & 1, ord = (f.packed >> 8) & 7;
			const int pi = B.perm[ord][i];
			if (((f.packed >> 12) & 15) == kBcNonconforming) {
				const S* y = a.zf + ((size_t)e * 4 + lf) * 3 * kNsNV * N + kNsNV * N;
#pragma unroll
				for (int k = 0; k < 2 * kNsNV; ++k) QF[slc][(lf * 2 * kNsNV + k) * N + i] = y[k * N + i];
				continue;
			}
			const S* own = TRo[slc] + lf * kNsNV * N;
			const S* nb = (TX[slc] + 4 * kNsNV * N) + lf * kNsNV * N;
			const typename NsArgs<S>::G* nJ = a.nJ + ((size_t)f.f * N + i) * 2;
			const int dst = side == 0 ? i : pi;
#pragma unroll
			for (int v = 0; v < kNsNV; ++v) {
				const S sL = side == 0 ? own[v * N + i] : nb[v * N + i];
				const S sR = side == 0 ? nb[v * N + pi] : own[v * N + pi];
				const S dl = S(0.5) * (sR - sL);
				QF[slc][((lf * 2 + 0) * kNsNV + v) * N + dst] = nJ[0] * dl;
				QF[slc][((lf * 2 + 1) * kNsNV + v) * N + dst] = nJ[1] * dl;
			}
		}
	}
	__syncthreads();
	if (valid) {
		// face numerical flux, items in the left ordering (as the Euler kernel)
		for (int item = t; item < 4 * N; item += Np) {
			const int lf = item / N, i = item % N;
			const ElemFace fc = B.ef[slc][lf];
			if (((fc.packed >> 12) & 15) == kBcNonconforming) {
				const S* z = a.zf + ((size_t)e * 4 + lf) * 3 * kNsNV * N;
#pragma unroll
				for (int v = 0; v < kNsNV; ++v) RES[slc][(lf * kNsNV + v) * N + i] = z[v * N + i];
				continue;
			}
			const int side = (fc.packed >> 4) & 1, ord = (fc.packed >> 8) & 7, fn = fc.packed & 15;
			const int en = fc.en >= 0 ? fc.en : e;
			const int pi = B.perm[ord][i];
			const S* own = TRo[slc] + lf * kNsNV * N;
			const S* nb = (TX[slc] + 4 * kNsNV * N) + lf * kNsNV * N;
			// left / right: element, local face, index of node i in that element's face ordering
			const int eLf = side == 0 ? e : en, lfL = side == 0 ? lf : fn, iL = i;
			const int eRf = side == 0 ? en : e, lfR = side == 0 ? fn : lf, iR = pi;
			const S* tvL = a.gvt + ((size_t)eLf * 4 + lfL) * 2 * kNsNV * N;
			const S* tvR = a.gvt + ((size_t)eRf * 4 + lfR) * 2 * kNsNV * N;
			const S kL = a.kap[((size_t)eLf * 4 + lfL) * N + iL], kR = a.kap[((size_t)eRf * 4 + lfR) * N + iR];
			const typename NsArgs<S>::G* nJ = a.nJ + ((size_t)fc.f * N + i) * 2;
			S sL[kNsNV], sR[kNsNV], gL[2 * kNsNV], gR[2 * kNsNV];
			const bool wall = fc.en < 0;   // boundary face: the boundary gradient is the interior one
			const int bcf = (fc.packed >> 12) & 15;
			const bool nf_provided = bcf == DGX_BC_NOSLIP_ADIABATIC || bcf == DGX_BC_NOSLIP_DIABATIC;
#pragma unroll
			for (int v = 0; v < kNsNV; ++v) {
				sL[v] = side == 0 ? own[v * N + i] : nb[v * N + i];
				sR[v] = side == 0 ? nb[v * N + pi] : own[v * N + pi];
				const S dl = S(0.5) * (sR[v] - sL[v]);
#pragma unroll
				for (int d = 0; d < 2; ++d) {
					const S q = nJ[d] * dl;
					S tL = tvL[(d * kNsNV + v) * N + iL], tR = tvR[(d * kNsNV + v) * N + iR];
					if constexpr (MODE == kNsDiag) {
						// the neighbour's gradient trace held fixed (its derivative is for its own seeded dof)
						if (side == 0) tR = strip(tR);
						else tL = strip(tL);
					} else if constexpr (MODE == kNsOffd) {
						// the own gradient trace held fixed, the neighbour's kept on the seeded face only
						const bool keep = lf == (int)blockIdx.y;
						if (side == 0) {
							tL = strip(tL);
							if (!keep) tR = strip(tR);
						} else {
							tR = strip(tR);
							if (!keep) tL = strip(tL);
						}
					}
					gL[v + kNsNV * d] = tL + eta * kL * q;
					gR[v + kNsNV * d] = wall ? gL[v + kNsNV * d] : tR + eta * kR * q;
				}
			}
			const S Jf = tsqrt(nJ[0] * nJ[0] + nJ[1] * nJ[1]);
			const S Jfi = trcp(Jf);
			const S nu[2] = { nJ[0] * Jfi, nJ[1] * Jfi };
			const int dst = side == 0 ? i : pi;
			if constexpr (PERT) {
				// FP64 states c_e + delta and fluxes; the normal Euler flux of c_e is subtracted before rounding
				const double* c0 = B.C[slc];
				double cL[kNsNV], cR[kNsNV], gLd[2 * kNsNV], gRd[2 * kNsNV], nd[2], fdn[kNsNV], cn[kNsNV];
				double fvL[2 * kNsNV], fvR[2 * kNsNV], cu[2];
#pragma unroll
				for (int v = 0; v < kNsNV; ++v) {
					cL[v] = c0[v] + (double)sL[v];
					cR[v] = c0[v] + (double)sR[v];
				}
#pragma unroll
				for (int k = 0; k < 2 * kNsNV; ++k) {
					gLd[k] = (double)gL[k];
					gRd[k] = (double)gR[k];
				}
				const double Jd = sqrt((double)nJ[0] * nJ[0] + (double)nJ[1] * nJ[1]);
				nd[0] = (double)nJ[0] / Jd;
				nd[1] = (double)nJ[1] / Jd;
				if constexpr (FLUX == DGX_FLUX_LF) numflux_lf<double, 2>(cL, cR, nd, fdn);
				else numflux_roe<double, 2>(cL, cR, nd, fdn);
				const double ci = 1.0 / c0[0];
				cu[0] = c0[1] * ci;
				cu[1] = c0[2] * ci;
				const double pc = kGM1 * (c0[3] - 0.5 * c0[0] * (cu[0] * cu[0] + cu[1] * cu[1]));
				euler_normal_flux<double, 2>(c0, cu, pc, nd, cn);
				ns_viscous_flux2<double>(cL, gLd, ns_mu<SUTH, double>(cL, (double)a.mu, a.r_s), (double)a.Pr, fvL);
				ns_viscous_flux2<double>(cR, gRd, ns_mu<SUTH, double>(cR, (double)a.mu, a.r_s), (double)a.Pr, fvR);
#pragma unroll
				for (int v = 0; v < kNsNV; ++v) {
					const double vis = nd[0] * 0.5 * (fvL[0 + 2 * v] + fvR[0 + 2 * v]) + nd[1] * 0.5 * (fvL[1 + 2 * v] + fvR[1 + 2 * v]);
					// no-slip flux walls: the energy flux is the provided one (numerical_flux_central_T.c:222-230)
					const double tot = (nf_provided && v == kNsNV - 1) ? a.walls.nfE : fdn[v] + vis;
					const S r = S((tot - cn[v]) * Jd);
					RES[slc][(lf * kNsNV + v) * N + dst] = side == 0 ? r : -r;
				}
			} else {
				S nnf[kNsNV], fvL[2 * kNsNV], fvR[2 * kNsNV];
				if constexpr (FLUX == DGX_FLUX_LF) numflux_lf<S, 2>(sL, sR, nu, nnf);
				else numflux_roe<S, 2>(sL, sR, nu, nnf);
				ns_viscous_flux2<S>(sL, gL, ns_mu<SUTH, S>(sL, a.mu, a.r_s), a.Pr, fvL);
				ns_viscous_flux2<S>(sR, gR, ns_mu<SUTH, S>(sR, a.mu, a.r_s), a.Pr, fvR);
#pragma unroll
				for (int v = 0; v < kNsNV; ++v) {
					const S vis = nu[0] * S(0.5) * (fvL[0 + 2 * v] + fvR[0 + 2 * v]) + nu[1] * S(0.5) * (fvL[1 + 2 * v] + fvR[1 + 2 * v]);
					// no-slip flux walls: the energy flux is the provided one (numerical_flux_central_T.c:222-230)
					const S tot = (nf_provided && v == kNsNV - 1) ? S(a.walls.nfE) : nnf[v] + vis;
					const S r = tot * Jf;
					RES[slc][(lf * kNsNV + v) * N + dst] = side == 0 ? r : -r;
				}
			}
		}
	}
	__syncthreads();   // the face traces are dead: X takes their place
	if (valid) {
		// G = G_v + sum_faces G_f at this node, then the volume flux
		S s[kNsNV], g[2 * kNsNV], f[2 * kNsNV], fv[2 * kNsNV];
		ns_volume_gradient<S, N>(B, a, slc, t, e, g);
		const S wq = B.w[c0] * B.w[c1];
		const S Jq = a.detJ[(size_t)e * Np + t];
		const S wji = trcp(wq * Jq);
#pragma unroll
		for (int lf = 0; lf < 4; ++lf) {
			const int fd = lf >> 1, cn = fd == 0 ? c0 : c1, at = fd == 0 ? c1 : c0;
			const S lw = B.Lift[lf & 1][cn] * B.w[at] * wji;
#pragma unroll
			for (int d = 0; d < 2; ++d)
#pragma unroll
				for (int v = 0; v < kNsNV; ++v) g[v + kNsNV * d] += lw * QF[slc][((lf * 2 + d) * kNsNV + v) * N + at];
		}
		if constexpr (PERT) {
			// FP64 pointwise fluxes of c_e + V minus the Euler flux of c_e; the contractions stay in S
			double sd[kNsNV], cd[kNsNV], gd[2 * kNsNV], fd[2 * kNsNV], fcd[2 * kNsNV], fvd[2 * kNsNV];
#pragma unroll
			for (int v = 0; v < kNsNV; ++v) {
				cd[v] = B.C[slc][v];
				sd[v] = cd[v] + (double)B.V[slc][v * Np + t];
			}
#pragma unroll
			for (int k = 0; k < 2 * kNsNV; ++k) gd[k] = (double)g[k];
			euler_flux<double, 2>(sd, fd);
			euler_flux<double, 2>(cd, fcd);
			ns_viscous_flux2<double>(sd, gd, ns_mu<SUTH, double>(sd, (double)a.mu, a.r_s), (double)a.Pr, fvd);
#pragma unroll
			for (int k = 0; k < 2 * kNsNV; ++k) f[k] = S(fd[k] - fcd[k] + fvd[k]);
		} else {
#pragma unroll
			for (int v = 0; v < kNsNV; ++v) s[v] = B.V[slc][v * Np + t];
			euler_flux<S, 2>(s, f);
			ns_viscous_flux2<S>(s, g, ns_mu<SUTH, S>(s, a.mu, a.r_s), a.Pr, fv);
#pragma unroll
			for (int k = 0; k < 2 * kNsNV; ++k) f[k] += fv[k];
		}
		const typename NsArgs<S>::G* m = a.met + (size_t)e * 4 * Np;
#pragma unroll
		for (int d0 = 0; d0 < 2; ++d0)
#pragma unroll
			for (int v = 0; v < kNsNV; ++v) {
				// fr_d0 = sum_d1 m(d1, d0) f_d1, m[(i + 2 j)] = J dxi_j / dx_i
				const S fr = m[(0 + 2 * d0) * Np + t] * f[0 + 2 * v] + m[(1 + 2 * d0) * Np + t] * f[1 + 2 * v];
				X[slc][(d0 * kNsNV + v) * Np + t] = wq * fr;
			}
	}
	__syncthreads();
	S z[kNsNV];
	if (valid) {
		// weak divergence sum_d Dh_d^T (W fr_d), minus the face lifts, at the GL nodes
#pragma unroll
		for (int v = 0; v < kNsNV; ++v) {
			S acc = S(0.0);
#pragma unroll
			for (int mm = 0; mm < N; ++mm) {
				acc += B.DhT[c0 * N + mm] * X[slc][(0 * kNsNV + v) * Np + mm + N * c1];
				acc += B.DhT[c1 * N + mm] * X[slc][(1 * kNsNV + v) * Np + c0 + N * mm];
			}
#pragma unroll
			for (int lf = 0; lf < 4; ++lf) {
				const int fd = lf >> 1, cn = fd == 0 ? c0 : c1, at = fd == 0 ? c1 : c0;
				acc -= B.Lift[lf & 1][cn] * B.w[at] * RES[slc][(lf * kNsNV + v) * N + at];
			}
			if constexpr (MODE == kNsJvp) {
				// - s_e M_e v: M = Ibar^T W J Ibar, so - s_e W J (Ibar v) here, before Ibar^T (V holds Ibar u + i Ibar v)
				if (a.mass_scale)
					acc.d -= a.mass_scale[e] * (rv(B.w[c0]) * rv(B.w[c1]) * a.detJ[(size_t)e * Np + t]) * B.V[slc][v * Np + t].d;
			}
			z[v] = acc;
		}
	}
	__syncthreads();
	if (valid)
#pragma unroll
		for (int v = 0; v < kNsNV; ++v) B.V[slc][v * Np + t] = z[v];
	__syncthreads();
	// back to the GLL test functions: Ibar^T = I^T along s, then along r
	if (valid)
#pragma unroll
		for (int v = 0; v < kNsNV; ++v) {
			S acc = S(0.0);
#pragma unroll
			for (int mm = 0; mm < N; ++mm) acc += B.I[mm * N + c1] * B.V[slc][v * Np + c0 + N * mm];
			z[v] = acc;
		}
	__syncthreads();
	if (valid)
#pragma unroll
		for (int v = 0; v < kNsNV; ++v) B.V[slc][v * Np + t] = z[v];
	__syncthreads();
	if (valid)
#pragma unroll
		for (int v = 0; v < kNsNV; ++v) {
			S acc = S(0.0);
#pragma unroll
			for (int mm = 0; mm < N; ++mm) acc += B.I[mm * N + c0] * B.V[slc][v * Np + mm + N * c1];
			if constexpr (MODE == kNsJvp) a.out[((size_t)e * kNsNV + v) * Np + t] = acc.d;
			else if constexpr (MODE == kNsDiag)
				a.out[(size_t)e * (kNsNV * Np) * (kNsNV * Np) + (size_t)(v * Np + t) * (kNsNV * Np) + a.col] = acc.d;
			else if constexpr (MODE == kNsOffd) {
				constexpr int nb = kNsNV * Np;
				const size_t blk = ((size_t)e * 4 + blockIdx.y) * nb;
				const int row = v * Np + t;
				if (a.offd_transposed) a.out[(blk + a.col) * nb + row] = acc.d;
				else a.out[(blk + row) * nb + a.col] = acc.d;
			} else a.out[((size_t)e * kNsNV + v) * Np + t] = acc;
		}
}

// phase 0: both kernels; 1: the gradient traces only; 2: the residual only (variable-order meshes run the
// face kernel between orders in between)
template <typename S, int N, bool PERT, int MODE>
hipError_t run_ns(Ctx& c, const NsArgs<S>& a, int phase)
{
	constexpr int EPB = ns_epb<S, N>(), Np = N * N;
	const int grid = (c.E + EPB - 1) / EPB;
	// kNsOffd: every element's gradient traces seeded at its own dof col (the kNsDiag gradient kernel), then one
	// residual launch per local face (blockIdx.y) seeded through that face's neighbour
	constexpr int GVMODE = (MODE == kNsOffd) ? kNsDiag : MODE;
	const dim3 rgrid(grid, MODE == kNsOffd ? 4 : 1);
	if (phase != 2) {
		hipLaunchKernelGGL((k_ns_gv<S, N, PERT, GVMODE>), dim3(grid), dim3(EPB * Np), 0, c.stream, a);
		DGX_HIP_CHECK(hipGetLastError());
	}
	if (phase == 1) return hipSuccess;
	const bool lf = c.cfg.num_flux == DGX_FLUX_LF;
	if (a.sutherland) {
		if (lf) hipLaunchKernelGGL((k_rhs_ns<S, N, DGX_FLUX_LF, PERT, MODE, true>), rgrid, dim3(EPB * Np), 0, c.stream, a);
		else hipLaunchKernelGGL((k_rhs_ns<S, N, DGX_FLUX_ROE, PERT, MODE, true>), rgrid, dim3(EPB * Np), 0, c.stream, a);
	} else {
		if (lf) hipLaunchKernelGGL((k_rhs_ns<S, N, DGX_FLUX_LF, PERT, MODE, false>), rgrid, dim3(EPB * Np), 0, c.stream, a);
		else hipLaunchKernelGGL((k_rhs_ns<S, N, DGX_FLUX_ROE, PERT, MODE, false>), rgrid, dim3(EPB * Np), 0, c.stream, a);
	}
	return hipGetLastError();
}

}  // namespace

template <typename S, bool PERT = false, int MODE = kNsRhs>
static hipError_t launch_ns(Ctx& c, const typename NsArgs<S>::G* u, typename NsArgs<S>::G* r,
                            const typename NsArgs<S>::G* met, const typename NsArgs<S>::G* detJ,
                            const typename NsArgs<S>::G* nJ, const Ops1D<typename NsArgs<S>::G>* ops,
                            const double* u64 = nullptr, int phase = 0, const double* v = nullptr, int col = 0,
                            const double* mass_scale = nullptr, int offd_transposed = 0)
{
	if (c.d != 2 || c.cfg.pde != DGX_PDE_NAVIER_STOKES || !c.ns_gvt || !c.ns_kappa) return hipErrorInvalidValue;
	if (c.E == 0) return hipSuccess;
	NsArgs<S> a{};
	a.u = u;
	a.u64 = u64;
	a.out = r;
	if constexpr (std::is_same<S, Dual>::value || std::is_same<S, Dual32>::value) {
		if constexpr (std::is_same<S, Dual>::value) {
			// FP64 dual numbers: their own gradient-trace buffers, allocated on first use
			const size_t ng = (size_t)c.E * 4 * 2 * c.nv * c.Nfc, nk = (size_t)c.E * 4 * c.Nfc;
			if (!c.ns_gvt_dual) DGX_HIP_CHECK(hipMalloc(&c.ns_gvt_dual, sizeof(S) * ng));
			if (!c.ns_kappa_dual) DGX_HIP_CHECK(hipMalloc(&c.ns_kappa_dual, sizeof(S) * nk));
			a.gvt = reinterpret_cast<S*>(c.ns_gvt_dual);
			a.kap = reinterpret_cast<S*>(c.ns_kappa_dual);
		} else {   // FP32 dual numbers are 8 bytes: the FP64-sized buffers hold them
			a.gvt = reinterpret_cast<S*>(c.ns_gvt);
			a.kap = reinterpret_cast<S*>(c.ns_kappa);
		}
		a.v = v;
		a.col = col;
		a.mass_scale = mass_scale;
		a.offd_transposed = offd_transposed;
	} else {
		a.gvt = reinterpret_cast<S*>(c.ns_gvt);     // FP32 uses the first half of the FP64-sized buffers
		a.kap = reinterpret_cast<S*>(c.ns_kappa);
	}
	a.met = met;
	a.detJ = detJ;
	a.nJ = nJ;
	a.ef = c.elem_faces;
	a.ops = ops;
	a.mu = S(c.cfg.mu);
	a.Pr = S(c.cfg.Pr);
	a.sutherland = c.cfg.viscosity_type == DGX_VISC_SUTHERLAND;
	a.r_s = c.cfg.r_s > 0.0 ? c.cfg.r_s : 1.0;
	a.E = c.E;
	if constexpr (std::is_same<S, float>::value) a.zf = c.hp_z32;
	else if constexpr (std::is_same<S, double>::value) a.zf = c.hp_z64;
	else a.zf = nullptr;   // dual numbers: no faces between orders
	for (int k = 0; k < 5; ++k) a.bc_prim[k] = c.cfg.bc_prim[k];
	a.walls = NsWalls{ c.cfg.wall_nf_E, c.cfg.wall_omega, c.cfg.wall_rho, c.cfg.wall_E };
	a.wall_xy = c.wall_xy;
	if (c.has_nonconforming && (!a.zf || MODE != kNsRhs)) return hipErrorInvalidValue;   // no linearisation there
	switch (c.N) {
	case 2: return run_ns<S, 2, PERT, MODE>(c, a, phase);
	case 3: return run_ns<S, 3, PERT, MODE>(c, a, phase);
	case 4: return run_ns<S, 4, PERT, MODE>(c, a, phase);
	case 5: return run_ns<S, 5, PERT, MODE>(c, a, phase);
	case 6: return run_ns<S, 6, PERT, MODE>(c, a, phase);
	default: return hipErrorInvalidValue;
	}
}

hipError_t launch_rhs_ns(Ctx& c, const double* u, double* r)
{
	return launch_ns<double>(c, u, r, c.metrics_vc, c.detJ_vc, c.nJ_fc, c.ops64);
}

hipError_t launch_rhs_ns32(Ctx& c, const float* u, float* r)
{
	return launch_ns<float>(c, u, r, c.metrics_vc32, c.detJ_vc32, c.nJ_fc32, c.ops32);
}

hipError_t launch_rhs_ns_phase(Ctx& c, const double* u, double* r, int phase)
{
	return launch_ns<double>(c, u, r, c.metrics_vc, c.detJ_vc, c.nJ_fc, c.ops64, nullptr, phase);
}

hipError_t launch_rhs_ns32_phase(Ctx& c, const float* u, float* r, int phase)
{
	return launch_ns<float>(c, u, r, c.metrics_vc32, c.detJ_vc32, c.nJ_fc32, c.ops32, nullptr, phase);
}

hipError_t launch_rhs_ns_mixed_phase(Ctx& c, const double* u, float* r, int phase)
{
	return launch_ns<float, true>(c, nullptr, r, c.metrics_vc32, c.detJ_vc32, c.nJ_fc32, c.ops32, u, phase);
}

hipError_t launch_rhs_ns_mixed(Ctx& c, const double* u, float* r)
{
	return launch_ns<float, true>(c, nullptr, r, c.metrics_vc32, c.detJ_vc32, c.nJ_fc32, c.ops32, u);
}

// the Navier-Stokes linearisation (the BR2 gradient differentiated with the rest, compute_rlhs_dg's 'i' branch):
// J v (with mass_scale: J v - s_e M_e v), and one column of every diagonal block
hipError_t launch_ns_jvp(Ctx& c, const double* u, const double* v, double* jv, const double* mass_scale)
{
	return launch_ns<Dual, false, kNsJvp>(c, u, jv, c.metrics_vc, c.detJ_vc, c.nJ_fc, c.ops64, nullptr, 0, v, 0,
	                                      mass_scale);
}

// the same product in FP32 dual numbers (FP64 state, direction, geometry and result): the FP32 inner operator of
// the implicit solve
hipError_t launch_ns_jvp32(Ctx& c, const double* u, const double* v, double* jv, const double* mass_scale)
{
	return launch_ns<Dual32, false, kNsJvp>(c, u, jv, c.metrics_vc, c.detJ_vc, c.nJ_fc, c.ops64, nullptr, 0, v, 0,
	                                        mass_scale);
}

hipError_t launch_ns_diag_col(Ctx& c, const double* u, double* blocks, int col)
{
	return launch_ns<Dual, false, kNsDiag>(c, u, blocks, c.metrics_vc, c.detJ_vc, c.nJ_fc, c.ops64, nullptr, 0,
	                                       nullptr, col);
}

// column col of every face block dR_e/du_n (E x 4 x nb x nb; the BR2 residual couples the neighbour's whole
// state through its lifted gradient, so the blocks are full), transposed ([e][lf][col][row]) or not
hipError_t launch_ns_offd_col(Ctx& c, const double* u, double* blocks, int col, bool transposed)
{
	return launch_ns<Dual, false, kNsOffd>(c, u, blocks, c.metrics_vc, c.detJ_vc, c.nJ_fc, c.ops64, nullptr, 0,
	                                       nullptr, col, nullptr, transposed ? 1 : 0);
}

}  // namespace dgx
