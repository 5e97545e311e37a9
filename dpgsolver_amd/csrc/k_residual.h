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
__device__ inline Dual32 operator*(float a, Dual32 b) { return Dual32(a * b.v, a * b.d); }

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
// MODE_VOL / MODE_FACE: the volume term alone (compute_volume_rlhs_dg) or the face term alone
// (compute_face_rlhs_dg), added into out (the reference accumulates both into a pre-zeroed rhs).
enum { MODE_JVP = 5, MODE_DIAG = 6, MODE_PERT = 7, MODE_OFFD = 8, MODE_VOL = 9, MODE_FACE = 10 };

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
	int grid;             // workgroups of the launch: 0 = one per element block, else a persistent grid
	unsigned long long* dmax;
	const double* mass_scale;   // MODE_JVP: out = J v - mass_scale[e] M_e v when not NULL
	const Real* zf;       // 2-D variable order: lifted fluxes of the faces to elements of another order
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

// Elements per workgroup: as many as fit 256 lanes and 56 KB of LDS, except HEX p=4 with float storage
// (the FP32 and mixed residuals), where one element per workgroup measured 5% faster on 64^3 elements
// (FP32 1.35 against 1.40 ms, mixed 1.73 against 1.82 ms, identical results); FP64 keeps two (2.45
// against 2.74 ms).
template <typename S, int D, int N, int NV>
constexpr int rhs_epb()
{
	if (D == 3 && N == 5 && sizeof(S) == 4) return 1;
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
#define DGX_RHS_PERSIST 0
#include "k_rhs_body.inc"
#undef DGX_RHS_PERSIST
}

// The same residual over a persistent grid (a.grid workgroups), each workgroup striding over its XCD's
// range of element blocks.  It holds values across the block loop and needs more registers (HEX p=4 FP32:
// about 120 VGPRs against 77), so the autotuner keeps it only where it measures faster.
template <typename S, typename R, int D, int N, int PDE, int FLUX, int MODE>
__global__ void __launch_bounds__(256) k_rhs_persistent(RhsArgs<S, R> a)
{
#define DGX_RHS_PERSIST 1
#include "k_rhs_body.inc"
#undef DGX_RHS_PERSIST
}

template <typename S, typename R, int D, int N, int PDE, int FLUX, int MODE, bool PERSIST = false>
inline hipError_t run_rhs(const RhsArgs<S, R>& a, hipStream_t st)
{
	constexpr int Np = ipow(N, D), NV = (PDE == DGX_PDE_EULER) ? D + 2 : 1;
	constexpr int EPB = rhs_epb<S, D, N, NV>();
	const int grid = (((a.E + EPB - 1) / EPB) + 7) / 8 * 8;
	if (a.E == 0) return hipSuccess;
	if constexpr (PERSIST) {
		const int pgrid = (a.grid > 0 && a.grid < grid) ? (a.grid + 7) / 8 * 8 : grid;
		hipLaunchKernelGGL((k_rhs_persistent<S, R, D, N, PDE, FLUX, MODE>), dim3(pgrid), dim3(EPB * Np), 0, st, a);
	} else {
		hipLaunchKernelGGL((k_rhs<S, R, D, N, PDE, FLUX, MODE>), dim3(grid), dim3(EPB * Np), 0, st, a);
	}
	return hipGetLastError();
}

// runtime (d, p, pde, flux) -> template instantiation.  PERSIST selects k_rhs_persistent; it is
// instantiated only in k_rhs_persistent.hip, because compiling both kernels of a configuration into one
// code object changes the register allocation of k_rhs (HEX p=4 FP64: 128 -> 167 VGPRs).
template <typename S, typename R, int MODE, bool PERSIST = false>
hipError_t dispatch_rhs(int d, int p, int pde, int flux, const RhsArgs<S, R>& a, hipStream_t st)
{
#define DGX_CASE_EULER(DD, NN)                                                              \
	if (d == DD && p + 1 == NN && pde == DGX_PDE_EULER) {                                   \
		if (flux == DGX_FLUX_LF) return run_rhs<S, R, DD, NN, DGX_PDE_EULER, DGX_FLUX_LF, MODE, PERSIST>(a, st); \
		return run_rhs<S, R, DD, NN, DGX_PDE_EULER, DGX_FLUX_ROE, MODE, PERSIST>(a, st);             \
	}
#define DGX_CASE_ADV(DD, NN)                                                                \
	if (d == DD && p + 1 == NN && pde == DGX_PDE_ADVECTION)                                 \
		return run_rhs<S, R, DD, NN, DGX_PDE_ADVECTION, DGX_FLUX_UPWIND, MODE, PERSIST>(a, st);
	DGX_CASE_EULER(2, 2) DGX_CASE_EULER(2, 3) DGX_CASE_EULER(2, 4) DGX_CASE_EULER(2, 5) DGX_CASE_EULER(2, 6)
	DGX_CASE_EULER(3, 2) DGX_CASE_EULER(3, 3) DGX_CASE_EULER(3, 4) DGX_CASE_EULER(3, 5)
	if constexpr (MODE != MODE_PERT) {
		DGX_CASE_ADV(2, 2) DGX_CASE_ADV(2, 3) DGX_CASE_ADV(2, 4) DGX_CASE_ADV(2, 5) DGX_CASE_ADV(2, 6)
	}
#undef DGX_CASE_EULER
#undef DGX_CASE_ADV
	return hipErrorInvalidValue;
}

// persistent-grid launches (defined in k_rhs_persistent.hip)
hipError_t dispatch_rhs_persistent(const Ctx& c, const RhsArgs<double, double>& a, int mode);
hipError_t dispatch_rhs_persistent(const Ctx& c, const RhsArgs<float, float>& a, int mode);

}  // namespace dgx
