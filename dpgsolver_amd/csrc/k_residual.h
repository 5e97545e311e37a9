// k_residual.h -- the fused element-centric DG residual kernel for QUAD/HEX (gfx950).
//
// One pass per element computes the whole rhs of compute_rlhs_common_dg
// (src/simulation/solvers/dg/solve_dg.c:182-189):
//   volume  rhs += sum_d tw1_vt_vc[d] fr_d          compute_volume_rlhs_T.c:145-177
//   faces   rhs -= tw0_vt_fc (J_f * nnf)             compute_face_rlhs_T.c:297-359
// but sum-factorised: with Ibar = I (x) I (x) I the GLL->GL interpolation and
// Dhat the GL-collocated derivative, tw1_d = Ibar^T Dhat_d^T W, so
//   rhs = Ibar^T [ sum_d Dhat_d^T (W fr_d) - sum_faces (Lift_side (x) W_f nnf) ],
//   Lift_side = I^{-T} e_side,
// and M^-1 rhs = Ibar^{-1} [ ... ] / (W detJ) for the explicit path (mass inverse
// for free).  Every face flux is evaluated by both neighbours from bitwise identical
// inputs (left trace, right trace permuted to the left ordering, left n*J_f), the way
// the reference evaluates it once per face, so the face coupling stays conservative.
//
// Data layout in HBM (reference layout, per element contiguous):
//   u, rhs     [E][nv][Np]               (sol_coef (Np, nv) column-major)
//   metrics    [E][d*d][Nc]              (metrics_vc (Nc, d, d))
//   nJ         [F][Nfc][d]               (normals_fc * jacobian_det_fc, left element)
//   elem faces [E][2d] {f, en, lfN | side<<4 | ord<<8 | bc<<12}
// Work mapping: one thread per volume node, EPB elements per workgroup, 1-D operators
// and face orientation tables in LDS, pencil sweeps through LDS (a thread owns one
// N-long line, reads it to registers, writes its N outputs).
#pragma once

#include "dgx_device.h"
#include "dgx_internal.h"

namespace dgx {

__device__ inline Dual operator*(double a, Dual b) { return Dual(a * b.v, a * b.d); }

// MODE_PERT: FP32 residual of an FP64 state in per-element perturbation form.  With c_e the state at
// the element's first node, the kernel interpolates and contracts u - c_e and F(u) - F(c_e) in FP32
// while every pointwise flux (volume and numerical) is evaluated in FP64 from c_e + (u - c_e).
// The constant part F(c_e) drops out exactly: its volume and face contributions cancel per test
// function (free-stream preservation), so the result is the full residual, and the FP32 rounding
// acts only on the small within-element variation instead of on the whole flux.
// MODE_OFFD: column `col` of the face (off-diagonal) blocks dR_e/du_n for every neighbour n.  The
// neighbour's state enters R_e only through its trace on the shared face, and the face integral tests
// only with the element's own face nodes (GLL nodal basis), so each block is dense on NFV = nv * N^(d-1)
// face-node rows and columns and zero elsewhere.  The seed sits on the neighbour's face node
// col = v * NF + i (i in the neighbour's face ordering face_vol_node(fn, i)); the element's own state
// carries none, so only the face flux of the face shared with n has a derivative; the faces are lifted
// one at a time and the face-node rows rho = v * NF + a (node face_vol_node(lf, a)) are stored
// transposed: out[((e * 2d + lf) * NFV + col) * NFV + rho].
enum { MODE_JVP = 5, MODE_DIAG = 6, MODE_PERT = 7, MODE_OFFD = 8 };

template <typename S, typename R>
struct RhsArgs {
	using Real = typename RealOf<S>::type;
	const Real* u;        // state
	const double* u64;    // FP64 state (MODE_PERT)
	const Real* v;        // JVP direction
	Real* out;            // residual (or diagonal blocks in MODE_DIAG)
	const R* met;
	const R* detJ;
	const R* nJ;
	const ElemFace* ef;
	const int* order;     // processing order (Morton order of element centroids)
	const Ops1D<R>* ops;
	double bc_prim[5];
	double b_adv[3];
	int E;
	int col;              // MODE_DIAG column (local dof)
	unsigned long long* dmax;
	const double* mass_scale;   // MODE_JVP: out = J v - mass_scale[e] M_e v when not NULL
};

template <int D, int N>
__device__ inline int face_vol_node(int lf, int i)
{
	const int fd = lf >> 1, fix = (lf & 1) ? N - 1 : 0;
	if (D == 2) return fd == 0 ? fix + N * i : i + N * fix;
	const int a = i % N, b = i / N;
	if (fd == 0) return fix + N * (a + N * b);
	if (fd == 1) return a + N * (fix + N * b);
	return a + N * (b + N * fix);
}

// In-place sweep along `dir` over NB blocks of N^DD values.
template <typename S, typename R, int DD, int N, int NB, int NT>
__device__ inline void sweep_inplace(S* buf, const R* M, int dir, int t)
{
	constexpr int BL = ipow(N, DD);
	constexpr int NL = ipow(N, DD - 1);
	constexpr int NPEN = NB * NL;
	constexpr int PPT = (NPEN + NT - 1) / NT;
	const int Sd = (dir == 0) ? 1 : (dir == 1 ? N : N * N);
	S in[PPT][N];
	int base[PPT];
#pragma unroll
	for (int k = 0; k < PPT; ++k) {
		const int pen = t + k * NT;
		base[k] = -1;
		if (pen < NPEN) {
			const int b = pen / NL, l = pen - b * NL;
			base[k] = b * BL + (l % Sd) + (l / Sd) * Sd * N;
#pragma unroll
			for (int q = 0; q < N; ++q) in[k][q] = buf[base[k] + q * Sd];
		}
	}
	__syncthreads();
#pragma unroll
	for (int k = 0; k < PPT; ++k) {
		if (base[k] >= 0) {
#pragma unroll
			for (int q = 0; q < N; ++q) {
				S acc = M[q * N] * in[k][0];
#pragma unroll
				for (int j = 1; j < N; ++j) acc += M[q * N + j] * in[k][j];
				buf[base[k] + q * Sd] = acc;
			}
		}
	}
	__syncthreads();
}

// Two independent in-place sweeps over one buffer: region A (NB1 blocks of N^D1 at offset 0) and
// region B (NB2 blocks of N^D2 at offset OFFB), sharing one read / barrier / write / barrier cycle.
template <typename S, typename R, int D1, int N, int NB1, int D2, int NB2, int OFFB, int NT>
__device__ inline void sweep_inplace2(S* buf, const R* M, int dir, int t)
{
	constexpr int BL1 = ipow(N, D1), NL1 = ipow(N, D1 - 1), NP1 = NB1 * NL1;
	constexpr int BL2 = ipow(N, D2), NL2 = ipow(N, D2 - 1), NP2 = NB2 * NL2;
	constexpr int NPEN = NP1 + NP2;
	constexpr int PPT = (NPEN + NT - 1) / NT;
	const int Sd = (dir == 0) ? 1 : (dir == 1 ? N : N * N);
	S in[PPT][N];
	int base[PPT];
#pragma unroll
	for (int k = 0; k < PPT; ++k) {
		const int pen = t + k * NT;
		int bs = -1;
		if (pen < NP1) {
			const int bb = pen / NL1, l = pen - bb * NL1;
			bs = bb * BL1 + (l % Sd) + (l / Sd) * Sd * N;
		} else if (pen < NPEN) {
			const int p2 = pen - NP1;
			const int bb = p2 / NL2, l = p2 - bb * NL2;
			bs = OFFB + bb * BL2 + (l % Sd) + (l / Sd) * Sd * N;
		}
		base[k] = bs;
		if (bs >= 0) {
#pragma unroll
			for (int q = 0; q < N; ++q) in[k][q] = buf[bs + q * Sd];
		}
	}
	__syncthreads();
#pragma unroll
	for (int k = 0; k < PPT; ++k) {
		if (base[k] >= 0) {
#pragma unroll
			for (int q = 0; q < N; ++q) {
				S acc = M[q * N] * in[k][0];
#pragma unroll
				for (int j = 1; j < N; ++j) acc += M[q * N + j] * in[k][j];
				buf[base[k] + q * Sd] = acc;
			}
		}
	}
	__syncthreads();
}

// 2-D interpolation (I along a, then along b) of NB blocks of N x N values in place, one block per
// thread (the last NB threads of the element), no barrier needed: each block is private to its thread.
template <typename S, typename R, int N, int NB, int NT>
__device__ inline void face_block_interp(S* buf, const R* M, int t)
{
	const int bi = t - (NT - NB);
	if (bi < 0) return;
	S* blk = buf + bi * N * N;
	S T[N][N];  // T[q][b] = sum_j M[q][j] blk[j + N b]
#pragma unroll
	for (int b = 0; b < N; ++b) {
		S row[N];
#pragma unroll
		for (int j = 0; j < N; ++j) row[j] = blk[j + N * b];
#pragma unroll
		for (int q = 0; q < N; ++q) {
			S acc = M[q * N] * row[0];
#pragma unroll
			for (int j = 1; j < N; ++j) acc += M[q * N + j] * row[j];
			T[q][b] = acc;
		}
	}
#pragma unroll
	for (int p = 0; p < N; ++p)
#pragma unroll
		for (int q = 0; q < N; ++q) {
			S acc = M[p * N] * T[q][0];
#pragma unroll
			for (int k = 1; k < N; ++k) acc += M[p * N + k] * T[q][k];
			blk[q + N * p] = acc;
		}
}

// dst (+)= sweep(src) along dir over NB blocks of N^DD values (src != dst).
template <typename S, typename R, int DD, int N, int NB, int NT, bool ACC>
__device__ inline void sweep_to(const S* src, S* dst, const R* M, int dir, int t)
{
	constexpr int BL = ipow(N, DD);
	constexpr int NL = ipow(N, DD - 1);
	constexpr int NPEN = NB * NL;
	constexpr int PPT = (NPEN + NT - 1) / NT;
	const int Sd = (dir == 0) ? 1 : (dir == 1 ? N : N * N);
#pragma unroll
	for (int k = 0; k < PPT; ++k) {
		const int pen = t + k * NT;
		if (pen < NPEN) {
			const int b = pen / NL, l = pen - b * NL;
			const int base = b * BL + (l % Sd) + (l / Sd) * Sd * N;
			S in[N];
#pragma unroll
			for (int q = 0; q < N; ++q) in[q] = src[base + q * Sd];
#pragma unroll
			for (int q = 0; q < N; ++q) {
				S acc = M[q * N] * in[0];
#pragma unroll
				for (int j = 1; j < N; ++j) acc += M[q * N + j] * in[j];
				if (ACC) dst[base + q * Sd] += acc;
				else dst[base + q * Sd] = acc;
			}
		}
	}
	__syncthreads();
}

// One pencil per thread (NB * N^(DD-1) <= NT): the pencil's base offset for each direction is computed once
// per thread and reused by every sweep of the kernel.
template <int DD, int N, int NB>
__device__ inline int pencil_base(int pen, int dir)
{
	constexpr int BL = ipow(N, DD), NL = ipow(N, DD - 1);
	if (pen >= NB * NL) return -1;
	const int Sd = (dir == 0) ? 1 : (dir == 1 ? N : N * N);
	const int b = pen / NL, l = pen - b * NL;
	return b * BL + (l % Sd) + (l / Sd) * Sd * N;
}

// in-place sweep along DIR of the thread's pencil (base < 0: idle thread, still takes both barriers)
template <typename S, typename R, int N, int DIR>
__device__ inline void sweep_inplace_pb(S* buf, const R* M, int base)
{
	constexpr int Sd = (DIR == 0) ? 1 : (DIR == 1 ? N : N * N);
	S in[N];
	if (base >= 0) {
#pragma unroll
		for (int q = 0; q < N; ++q) in[q] = buf[base + q * Sd];
	}
	__syncthreads();
	if (base >= 0) {
#pragma unroll
		for (int q = 0; q < N; ++q) {
			S acc = M[q * N] * in[0];
#pragma unroll
			for (int j = 1; j < N; ++j) acc += M[q * N + j] * in[j];
			buf[base + q * Sd] = acc;
		}
	}
	__syncthreads();
}

// dst (+)= sweep(src) along DIR for the thread's pencil (src != dst); the accumulation starts from dst
template <typename S, typename R, int N, int DIR, bool ACC>
__device__ inline void sweep_to_pb(const S* src, S* dst, const R* M, int base)
{
	constexpr int Sd = (DIR == 0) ? 1 : (DIR == 1 ? N : N * N);
	if (base >= 0) {
		S in[N];
#pragma unroll
		for (int q = 0; q < N; ++q) in[q] = src[base + q * Sd];
#pragma unroll
		for (int q = 0; q < N; ++q) {
			S acc;
			if (ACC) {
				acc = dst[base + q * Sd];
				acc += M[q * N] * in[0];
			} else {
				acc = M[q * N] * in[0];
			}
#pragma unroll
			for (int j = 1; j < N; ++j) acc += M[q * N + j] * in[j];
			dst[base + q * Sd] = acc;
		}
	}
	__syncthreads();
}

template <typename S, int D, int N, int NV>
constexpr int rhs_epb()
{
	constexpr int Np = ipow(N, D), NF = ipow(N, D - 1), NFACE = 2 * D;
	constexpr int trsz = NFACE * NV * NF + (NFACE * NF > Np ? NFACE * NF : Np) * NV;
	constexpr int bytes = (int)sizeof(S) * (NV * Np + trsz) + NFACE * 16;
	constexpr int by_lds = (56 * 1024) / bytes;
	constexpr int by_thr = 256 / Np;
	constexpr int e = by_lds < by_thr ? by_lds : by_thr;
	return e > 0 ? e : 1;
}

template <typename S, typename R, int D, int N, int PDE, int FLUX, int MODE>
__global__ void __launch_bounds__(256) k_rhs(RhsArgs<S, R> a)
{
	constexpr int Np = ipow(N, D), NF = ipow(N, D - 1), NV = (PDE == DGX_PDE_EULER) ? D + 2 : 1;
	constexpr int NFACE = 2 * D, NT = Np;
	constexpr int EPB = rhs_epb<S, D, N, NV>();
	constexpr bool MINV = (MODE == MODE_RHS_MINV);

	__shared__ R sI[N * N], sIT[N * N], sDhT[N * N], sIinv[N * N], sLift[2][N], sW[N];
	__shared__ int sPerm[8][NF];
	// traces: own half, then the neighbour half which is reused as the flux staging buffer
	constexpr int TRSZ = NFACE * NV * NF + (NFACE * NF > Np ? NFACE * NF : Np) * NV;
	__shared__ S sBuf[EPB][NV * Np + TRSZ];   // volume block U, then the face traces TR
	__shared__ ElemFace sEF[EPB][NFACE];
	__shared__ unsigned long long sMax;
	constexpr bool PERT = (MODE == MODE_PERT);
	__shared__ double sC[EPB][PERT ? NV : 1];   // MODE_PERT reference state c_e

	// XCD-aware schedule: the grid is a multiple of 8 and consecutive workgroups land on different
	// XCDs, so XCD x takes the x-th contiguous range of the (Morton-ordered) element sequence and
	// neighbour traces are read from its own L2.
	const int per_xcd = gridDim.x >> 3;
	const int lblk = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
	if (lblk * EPB >= a.E) return;  // whole workgroup idle (padding), before any barrier
	const int slot = threadIdx.x / NT, t = threadIdx.x - slot * NT;
	const int pos = lblk * EPB + slot;
	const bool valid = (slot < EPB) && (pos < a.E);
	const int e = valid ? a.order[pos] : 0;
	const int sl = slot < EPB ? slot : 0;

	for (int i = threadIdx.x; i < N * N; i += blockDim.x) {
		sI[i] = a.ops->I[i];
		sIT[i] = a.ops->IT[i];
		sDhT[i] = a.ops->DhT[i];
		sIinv[i] = a.ops->Iinv[i];
	}
	for (int i = threadIdx.x; i < N; i += blockDim.x) {
		sLift[0][i] = a.ops->Lift[0][i];
		sLift[1][i] = a.ops->Lift[1][i];
		sW[i] = a.ops->w[i];
	}
	for (int i = threadIdx.x; i < 8 * NF; i += blockDim.x) sPerm[i / NF][i % NF] = a.ops->perm[i / NF][i % NF];
	if (threadIdx.x == 0) sMax = 0ull;
	if (valid && t < NFACE) sEF[sl][t] = a.ef[(size_t)e * NFACE + t];

	S* U = sBuf[sl];
	S* TR = U + NV * Np;
	S* BUF = TR + NFACE * NV * NF;   // neighbour-trace half, reused for the flux components
	constexpr bool PB = (NV * ipow(N, D - 1) <= NT);   // one volume pencil per thread
	int pb[3] = { -1, -1, -1 };
	if constexpr (PB) {
#pragma unroll
		for (int dir = 0; dir < D; ++dir) pb[dir] = pencil_base<D, N, NV>(t, dir);
	}
	const size_t ebase = (size_t)e * NV * Np;
	if (valid) {
		if constexpr (PERT) {
			if (t < NV) sC[sl][t] = a.u64[ebase + (size_t)t * Np];
		}
		for (int i = t; i < NV * Np; i += NT) {
			if constexpr (PERT) U[i] = S(a.u64[ebase + i] - a.u64[ebase + (size_t)(i / Np) * Np]);
			else if constexpr (MODE == MODE_JVP) U[i] = S(a.u[ebase + i], a.v[ebase + i]);
			else if constexpr (MODE == MODE_DIAG) U[i] = S(a.u[ebase + i], i == a.col ? 1.0 : 0.0);
			else U[i] = S(a.u[ebase + i]);
		}
	}
	__syncthreads();

	// ---- traces at the GLL face nodes: own (from LDS) and neighbour (gathered from HBM);
	//      one (face, node) item per thread, the variable loop reuses the address arithmetic
	if (valid) {
		for (int item = t; item < NFACE * NF; item += NT) {
			const int lf = item / NF, i = item - lf * NF;
			const int vown = face_vol_node<D, N>(lf, i);
			const ElemFace ef = sEF[sl][lf];
			S* to = TR + lf * NV * NF + i;
			S* tn = TR + (NFACE + lf) * NV * NF + i;
			if (ef.en >= 0) {
				const size_t gi = (size_t)ef.en * NV * Np + face_vol_node<D, N>(ef.packed & 15, i);
#pragma unroll
				for (int v = 0; v < NV; ++v) {
					to[v * NF] = U[v * Np + vown];
					if constexpr (PERT) tn[v * NF] = S(a.u64[gi + v * Np] - sC[sl][v]);
					else if constexpr (MODE == MODE_OFFD)
						tn[v * NF] = S(a.u[gi + v * Np], (v * NF + i == a.col) ? 1.0 : 0.0);
					else if constexpr (MODE == MODE_JVP) tn[v * NF] = S(a.u[gi + v * Np], a.v[gi + v * Np]);
					else tn[v * NF] = S(a.u[gi + v * Np]);
				}
			} else {
#pragma unroll
				for (int v = 0; v < NV; ++v) {
					to[v * NF] = U[v * Np + vown];
					tn[v * NF] = S(0.0);
				}
			}
		}
	}
	__syncthreads();  // traces written by other threads are read by the first sweep
	// ---- GLL -> GL: volume (D sweeps) and face traces (D-1 sweeps) sharing barrier cycles
	if constexpr (D == 3 && 2 * NFACE * NV <= NT) {
		// face traces: one N x N block per thread (in place, private), then the volume sweeps
		face_block_interp<S, R, N, 2 * NFACE * NV, NT>(TR, sI, t);
		if constexpr (PB) {
			sweep_inplace_pb<S, R, N, 0>(U, sI, pb[0]);
			sweep_inplace_pb<S, R, N, 1>(U, sI, pb[1]);
			sweep_inplace_pb<S, R, N, 2>(U, sI, pb[2]);
		} else {
#pragma unroll
			for (int dir = 0; dir < D; ++dir) sweep_inplace<S, R, D, N, NV, NT>(U, sI, dir, t);
		}
	} else {
#pragma unroll
		for (int dir = 0; dir < D; ++dir) {
			if (dir < D - 1) sweep_inplace2<S, R, D, N, NV, D - 1, 2 * NFACE * NV, NV * Np, NT>(U, sI, dir, t);
			else sweep_inplace<S, R, D, N, NV, NT>(U, sI, dir, t);
		}
	}

	// ---- numerical flux at the face cubature (canonical = left ordering)
	constexpr int NIT = NFACE * NF;
	constexpr int IPT = (NIT + NT - 1) / NT;
	S res[IPT][NV];
	int dst[IPT];
#pragma unroll
	for (int k = 0; k < IPT; ++k) {
		const int item = t + k * NT;
		dst[k] = -1;
		if (valid && item < NIT) {
			const int lf = item / NF, i = item - lf * NF;
			const ElemFace ef = sEF[sl][lf];
			const int side = (ef.packed >> 4) & 1, ord = (ef.packed >> 8) & 7, bc = (ef.packed >> 12) & 15;
			const S* own = TR + lf * NV * NF;
			const S* nbr = TR + (NFACE + lf) * NV * NF;
			const int pi = sPerm[ord][i];
			S sL[NV], sR[NV];
			if (side == 0) {
#pragma unroll
				for (int v = 0; v < NV; ++v) { sL[v] = own[v * NF + i]; sR[v] = nbr[v * NF + pi]; }
			} else {
#pragma unroll
				for (int v = 0; v < NV; ++v) { sL[v] = nbr[v * NF + i]; sR[v] = own[v * NF + pi]; }
			}
			const R* np_ = a.nJ + ((size_t)ef.f * NF + i) * D;
			R nJv[D];
#pragma unroll
			for (int k2 = 0; k2 < D; ++k2) nJv[k2] = np_[k2];
			R J2 = nJv[0] * nJv[0];
#pragma unroll
			for (int k2 = 1; k2 < D; ++k2) J2 += nJv[k2] * nJv[k2];
			const R Jf = tsqrt(J2);
			const R Jfi = trcp(Jf);
			S nu[D];
#pragma unroll
			for (int k2 = 0; k2 < D; ++k2) nu[k2] = S(nJv[k2] * Jfi);
			S nnf[NV];
			if constexpr (PERT && PDE == DGX_PDE_EULER) {
				// FP64 states c_e + delta, FP64 numerical flux minus the normal flux of c_e
				double cL[NV], cR[NV], nd[D], fd[NV], cu[D], cn[NV];
#pragma unroll
				for (int v = 0; v < NV; ++v) {
					cL[v] = sC[sl][v] + (double)sL[v];
					cR[v] = sC[sl][v] + (double)sR[v];
				}
#pragma unroll
				for (int k2 = 0; k2 < D; ++k2) nd[k2] = (double)nJv[k2] / (double)Jf;
				if (ef.en < 0) bc_state<double, D>(bc, cL, nd, a.bc_prim, cR);
				if constexpr (FLUX == DGX_FLUX_LF) numflux_lf<double, D>(cL, cR, nd, fd);
				else numflux_roe<double, D>(cL, cR, nd, fd);
				const double* c0 = sC[sl];
				const double ci = 1.0 / c0[0];
				double V2c = 0.0;
#pragma unroll
				for (int k2 = 0; k2 < D; ++k2) { cu[k2] = c0[1 + k2] * ci; V2c += cu[k2] * cu[k2]; }
				const double pc = kGM1 * (c0[D + 1] - 0.5 * c0[0] * V2c);
				euler_normal_flux<double, D>(c0, cu, pc, nd, cn);
#pragma unroll
				for (int v = 0; v < NV; ++v) nnf[v] = S(fd[v] - cn[v]);
			} else if constexpr (PDE == DGX_PDE_EULER) {
				if (ef.en < 0) bc_state<S, D>(bc, sL, nu, a.bc_prim, sR);
				if constexpr (FLUX == DGX_FLUX_LF) numflux_lf<S, D>(sL, sR, nu, nnf);
				else numflux_roe<S, D>(sL, sR, nu, nnf);
			} else {
				double bn = a.b_adv[0] * rv(nu[0]);
#pragma unroll
				for (int k2 = 1; k2 < D; ++k2) bn += a.b_adv[k2] * rv(nu[k2]);
				nnf[0] = (bn > kEps) ? R(bn) * sL[0] : R(bn) * sR[0];
			}
#pragma unroll
			for (int v = 0; v < NV; ++v) res[k][v] = (side == 0) ? Jf * nnf[v] : R(-1.0) * (Jf * nnf[v]);
			dst[k] = lf * NV * NF + (side == 0 ? i : pi);
		}
	}
	__syncthreads();
#pragma unroll
	for (int k = 0; k < IPT; ++k)
		if (dst[k] >= 0) {
#pragma unroll
			for (int v = 0; v < NV; ++v) TR[dst[k] + v * NF] = res[k][v];
		}

	// ---- physical flux at the volume cubature, contravariant and weighted, kept in registers
	S fr[D][NV];
	double vgl[(MODE == MODE_JVP) ? NV : 1];   // MODE_JVP: (Ibar v) at this node, for the mass term
	if (valid && MODE != MODE_OFFD) {
		const int q = t;
		S s[NV];
#pragma unroll
		for (int v = 0; v < NV; ++v) s[v] = U[v * Np + q];
		if constexpr (MODE == MODE_JVP) {
#pragma unroll
			for (int v = 0; v < NV; ++v) vgl[v] = s[v].d;
		}
		R m[D * D];
		const R* mp = a.met + (size_t)e * D * D * Np;
#pragma unroll
		for (int i = 0; i < D * D; ++i) m[i] = mp[i * Np + q];
		const int c0 = q % N, c1 = (q / N) % N, c2 = (D == 3) ? q / (N * N) : 0;
		R wq = sW[c0] * sW[c1];
		if (D == 3) wq = wq * sW[c2];
		if constexpr (PERT && PDE == DGX_PDE_EULER) {
			// physical fluxes of c_e + delta and of c_e in FP64, their difference contracted in FP32
			double cs[NV], fu[D * NV], fc[D * NV];
#pragma unroll
			for (int v = 0; v < NV; ++v) cs[v] = sC[sl][v] + (double)s[v];
			euler_flux<double, D>(cs, fu);
			euler_flux<double, D>(sC[sl], fc);
#pragma unroll
			for (int v = 0; v < NV; ++v) {
				S df[D];
#pragma unroll
				for (int k = 0; k < D; ++k) df[k] = S(fu[k + D * v] - fc[k + D * v]);
#pragma unroll
				for (int d0 = 0; d0 < D; ++d0) {
					S acc = m[0 + D * d0] * df[0];
#pragma unroll
					for (int k = 1; k < D; ++k) acc += m[k + D * d0] * df[k];
					fr[d0][v] = wq * acc;
				}
			}
		} else if constexpr (PDE == DGX_PDE_EULER) {
			euler_contravariant_flux<S, R, D>(s, m, wq, fr);
		} else {
#pragma unroll
			for (int d0 = 0; d0 < D; ++d0) {
				R bm = R(a.b_adv[0]) * m[0 + D * d0];
#pragma unroll
				for (int k = 1; k < D; ++k) bm += R(a.b_adv[k]) * m[k + D * d0];
				fr[d0][0] = wq * (bm * s[0]);
			}
		}
	}

	// ---- weak divergence in GL space: U = sum_d Dhat_d^T fr_d, one direction at a time through BUF
	//      (MODE_OFFD: the volume term has no derivative and is skipped)
#pragma unroll
	for (int d0 = 0; d0 < (MODE == MODE_OFFD ? 0 : D); ++d0) {
		if (valid) {
#pragma unroll
			for (int v = 0; v < NV; ++v) BUF[v * Np + t] = fr[d0][v];
		}
		__syncthreads();
		if constexpr (PB && D == 3) {
			if (d0 == 0) sweep_to_pb<S, R, N, 0, false>(BUF, U, sDhT, pb[0]);
			else if (d0 == 1) sweep_to_pb<S, R, N, 1, true>(BUF, U, sDhT, pb[1]);
			else sweep_to_pb<S, R, N, 2, true>(BUF, U, sDhT, pb[2]);
		} else {
			if (d0 == 0) sweep_to<S, R, D, N, NV, NT, false>(BUF, U, sDhT, 0, t);
			else sweep_to<S, R, D, N, NV, NT, true>(BUF, U, sDhT, d0, t);
		}
	}

	// ---- face lifts (and the mass inverse) at the GL nodes
	// per-face coefficient Lift_side[c_fd] * W_f and face-node offset, shared by all variables
	R cf[NFACE];
	int of[NFACE];
	R scale = R(1.0);
	if (valid) {
		const int q = t;
		const int cc[3] = { q % N, (q / N) % N, (D == 3) ? q / (N * N) : 0 };
		if constexpr (MINV) {
			R wq = sW[cc[0]] * sW[cc[1]];
			if (D == 3) wq = wq * sW[cc[2]];
			scale = trcp(wq * a.detJ[(size_t)e * Np + q]);
		}
#pragma unroll
		for (int lf = 0; lf < NFACE; ++lf) {
			const int fd = lf >> 1;
			int fa, fb;
			if (fd == 0) { fa = cc[1]; fb = cc[2]; }
			else if (fd == 1) { fa = cc[0]; fb = cc[2]; }
			else { fa = cc[0]; fb = cc[1]; }
			R wf = sW[fa];
			if (D == 3) wf = wf * sW[fb];
			cf[lf] = sLift[lf & 1][cc[fd]] * wf;
			of[lf] = lf * NV * NF + ((D == 3) ? fa + N * fb : fa);
		}
	}
	if constexpr (MODE == MODE_OFFD) {
		constexpr int NFV = NV * NF;
		for (int lf = 0; lf < NFACE; ++lf) {
			if (valid) {
#pragma unroll
				for (int v = 0; v < NV; ++v) U[v * Np + t] = S(0.0) - cf[lf] * TR[of[lf] + v * NF];
			}
			__syncthreads();
#pragma unroll
			for (int dir = 0; dir < D; ++dir) sweep_inplace<S, R, D, N, NV, NT>(U, sIT, dir, t);
			if (valid) {
				double* ob = a.out + (((size_t)e * NFACE + lf) * NFV + a.col) * NFV;
				for (int r = t; r < NFV; r += NT) {
					const int v = r / NF;
					ob[r] = U[v * Np + face_vol_node<D, N>(lf, r - v * NF)].d;
				}
			}
			__syncthreads();
		}
		return;
	}
	if (valid) {
		const int q = t;
		double msc = 0.0;   // MODE_JVP mass term: - s_e W_q J_q (Ibar v)_q before the final Ibar^T
		if constexpr (MODE == MODE_JVP) {
			if (a.mass_scale) {
				const int c0 = q % N, c1 = (q / N) % N, c2 = (D == 3) ? q / (N * N) : 0;
				double wq = sW[c0] * sW[c1];
				if (D == 3) wq *= sW[c2];
				msc = a.mass_scale[e] * wq * a.detJ[(size_t)e * Np + q];
			}
		}
#pragma unroll
		for (int v = 0; v < NV; ++v) {
			S acc = U[v * Np + q];
#pragma unroll
			for (int lf = 0; lf < NFACE; ++lf) acc -= cf[lf] * TR[of[lf] + v * NF];
			if constexpr (MINV) acc = scale * acc;
			if constexpr (MODE == MODE_JVP) acc.d -= msc * vgl[v];
			U[v * Np + q] = acc;
		}
	}
	__syncthreads();
	// ---- back to the GLL test functions: Ibar^T (rhs) or Ibar^{-1} (M^-1 rhs)
	if constexpr (PB && D == 3) {
		sweep_inplace_pb<S, R, N, 0>(U, MINV ? sIinv : sIT, pb[0]);
		sweep_inplace_pb<S, R, N, 1>(U, MINV ? sIinv : sIT, pb[1]);
		sweep_inplace_pb<S, R, N, 2>(U, MINV ? sIinv : sIT, pb[2]);
	} else {
#pragma unroll
		for (int dir = 0; dir < D; ++dir) sweep_inplace<S, R, D, N, NV, NT>(U, MINV ? sIinv : sIT, dir, t);
	}

	if (valid) {
		double lmax = 0.0;
		for (int i = t; i < NV * Np; i += NT) {
			if constexpr (MODE == MODE_JVP) a.out[ebase + i] = U[i].d;
			else if constexpr (MODE == MODE_DIAG)
				a.out[(size_t)e * (NV * Np) * (NV * Np) + (size_t)i * (NV * Np) + a.col] = U[i].d;
			else if constexpr (MODE == MODE_OFFD) {
				// stored per face above
			} else {
				a.out[ebase + i] = U[i];
				lmax = fmax(lmax, fabs((double)rv(U[i])));
			}
		}
		if constexpr (MODE == MODE_RHS_MINV) {
			if (a.dmax) atomicMax(&sMax, (unsigned long long)__double_as_longlong(lmax));
		}
	}
	if constexpr (MODE == MODE_RHS_MINV) {
		__syncthreads();
		if (threadIdx.x == 0 && a.dmax) atomicMax(a.dmax, sMax);
	}
}

template <typename S, typename R, int D, int N, int PDE, int FLUX, int MODE>
inline hipError_t run_rhs(const RhsArgs<S, R>& a, hipStream_t st)
{
	constexpr int Np = ipow(N, D), NV = (PDE == DGX_PDE_EULER) ? D + 2 : 1;
	constexpr int EPB = rhs_epb<S, D, N, NV>();
	const int grid = (((a.E + EPB - 1) / EPB) + 7) / 8 * 8;
	if (a.E == 0) return hipSuccess;
	hipLaunchKernelGGL((k_rhs<S, R, D, N, PDE, FLUX, MODE>), dim3(grid), dim3(EPB * Np), 0, st, a);
	return hipGetLastError();
}

// runtime (d, p, pde, flux) -> template instantiation
template <typename S, typename R, int MODE>
hipError_t dispatch_rhs(int d, int p, int pde, int flux, const RhsArgs<S, R>& a, hipStream_t st)
{
#define DGX_CASE_EULER(DD, NN)                                                              \
	if (d == DD && p + 1 == NN && pde == DGX_PDE_EULER) {                                   \
		if (flux == DGX_FLUX_LF) return run_rhs<S, R, DD, NN, DGX_PDE_EULER, DGX_FLUX_LF, MODE>(a, st); \
		return run_rhs<S, R, DD, NN, DGX_PDE_EULER, DGX_FLUX_ROE, MODE>(a, st);             \
	}
#define DGX_CASE_ADV(DD, NN)                                                                \
	if (d == DD && p + 1 == NN && pde == DGX_PDE_ADVECTION)                                 \
		return run_rhs<S, R, DD, NN, DGX_PDE_ADVECTION, DGX_FLUX_UPWIND, MODE>(a, st);
	DGX_CASE_EULER(2, 2) DGX_CASE_EULER(2, 3) DGX_CASE_EULER(2, 4) DGX_CASE_EULER(2, 5) DGX_CASE_EULER(2, 6)
	DGX_CASE_EULER(3, 2) DGX_CASE_EULER(3, 3) DGX_CASE_EULER(3, 4) DGX_CASE_EULER(3, 5)
	if constexpr (MODE != MODE_PERT) {
		DGX_CASE_ADV(2, 2) DGX_CASE_ADV(2, 3) DGX_CASE_ADV(2, 4) DGX_CASE_ADV(2, 5) DGX_CASE_ADV(2, 6)
	}
#undef DGX_CASE_EULER
#undef DGX_CASE_ADV
	return hipErrorInvalidValue;
}

}  // namespace dgx
