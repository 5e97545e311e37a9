// dgx_internal.h -- shared definitions of the MI355X DG residual library (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/dgx.h"

namespace dgx {

constexpr int kMaxN = 8;          // p <= 7
constexpr int kMaxNF = kMaxN * kMaxN;

constexpr double kGamma = 1.4;    // definitions_physics.h:23-25
constexpr double kGM1 = 0.4;
constexpr double kEps = 1.0e-15;  // definitions_tol.h:23
constexpr double kEpsPhys = 1.0e-13;

// 1-D operator tables for one polynomial order, row-major [out][in].
template <typename T>
struct Ops1D {
	T I[kMaxN * kMaxN];      // GLL -> GL interpolation
	T IT[kMaxN * kMaxN];     // transpose of I (test-function side)
	T DhT[kMaxN * kMaxN];    // transpose of the GL-collocated derivative
	T Iinv[kMaxN * kMaxN];   // inverse of I (GL -> GLL)
	T Dg[kMaxN * kMaxN];     // GLL-collocated derivative (metric nodes)
	T B[kMaxN * kMaxN];      // Bernstein basis at the GLL nodes (Bezier -> nodal)
	T Binv[kMaxN * kMaxN];   // nodal -> Bezier
	T Ext[2][kMaxN];         // GLL basis at r = -1 / +1 (unit vectors)
	T Lift[2][kMaxN];        // I^{-T} e_side : face lift expressed at the GL volume nodes
	T w[kMaxN];              // GL weights
	T wgll[kMaxN];           // GLL weights
	int perm[8][kMaxNF];     // face orientation permutations, indexed by ord_bits()
};

// The reference's face orientation index ind_ord at the ABI (the case of the face node correspondence,
// src/element/nodes_correspondence.c) -> the orientation bits that index Ops1D::perm (bit0 transpose, bit1
// reverse the first face coordinate, bit2 reverse the second); -1 for a code that means nothing in dimension d.
// LINE faces (d = 2): 0 aligned, 1 reversed.  QUAD faces (d = 3): cases 0-3 reverse the first and/or second
// coordinate, cases 4-7 are the transposed numbering, case 5 = transpose then reverse the first, case 6 =
// transpose then reverse the second.
inline int ord_bits(int d, int ind_ord)
{
	if (d == 1) return ind_ord == 0 ? 0 : -1;
	if (d == 2) return ind_ord == 0 ? 0 : (ind_ord == 1 ? 2 : -1);
	static const int kBits[8] = { 0, 2, 4, 6, 1, 3, 5, 7 };
	return (ind_ord >= 0 && ind_ord < 8) ? kBits[ind_ord] : -1;
}

// Variable-order 2-D meshes: orders 1..kHpMaxP, faces between elements of different orders.
constexpr int kHpMaxN = 6;
// bc code of an element's face to an element of another order (internal: the face kernel lifts its flux)
constexpr int kBcNonconforming = 15;

// 1-D operators between orders, N = p + 1 (index 0 and 1 unused):
// Ig[Ns][Nf][q * Ns + i]: GLL_Ns Lagrange basis i at GL_Nf point q (traces to the face cubature);
// Pg[Ne][Nf][q * Ne + a]: GL_Ne Lagrange basis a at GL_Nf point q (exact for the element's test functions);
// w / winv: GL weights of each order and their inverses.
template <typename T>
struct HpOps {
	T Ig[kHpMaxN + 1][kHpMaxN + 1][kHpMaxN * kHpMaxN];
	T Pg[kHpMaxN + 1][kHpMaxN + 1][kHpMaxN * kHpMaxN];
	T w[kHpMaxN + 1][kHpMaxN];
	T winv[kHpMaxN + 1][kHpMaxN];
};

// One face between elements of different orders: first state value of each side's element, first value of
// each side's lifted-flux block (own face, [var][own GL face point]), and NL | NR << 3 | lfL << 6 | lfR << 8 |
// ord << 10.
struct HpFace {
	long long uL, uR, zL, zR;
	int info;
	int eL, eR;          // each side's element index within its order group
	int pad;
};

// TRI elements of the reference's mixed 2-D meshes (dgx_tri.cpp): Lagrange basis on the WSH nodes of order p,
// WV cubature of degree 2p, GL face points of the element's own order.  Row-major [out][in] unless noted.
constexpr int kTriMaxP = 5;
constexpr int kTriMaxNp = 21;   // (p+1)(p+2)/2 at p = 5
constexpr int kTriMaxNc = 25;   // WV10
template <typename T>
struct TriOps {
	T I[kTriMaxNc * kTriMaxNp];        // basis at the cubature [q][m]
	T Dr[kTriMaxNc * kTriMaxNp];       // d/dr, d/ds of the basis at the cubature [q][m]
	T Ds[kTriMaxNc * kTriMaxNp];
	T DrW[kTriMaxNp * kTriMaxNc];      // transposed and weighted: [m][q] = d phi_m/dr (q) w_q
	T DsW[kTriMaxNp * kTriMaxNc];
	T w[kTriMaxNc];                    // cubature weights
	T Ifc[3][kHpMaxN * kTriMaxNp];     // basis at the element's own GL points of face lf [a][m]
	T tw0[3][kTriMaxNp * kHpMaxN];     // test functions times GL weights at those points [m][a]
	// the inverse mass matrix and the lift traces K_f = Ifc_f M^-1 tw0_f of an element with det J = 1 (an affine
	// element's are these divided by its det J)
	T Minv[kTriMaxNp * kTriMaxNp];
	T K[3][kHpMaxN * kHpMaxN];
	// the lifts of the three faces' projections y at the cubature, I M^-1 tw0_f side by side: [q][f N + b]
	T L[kTriMaxNc * 3 * kHpMaxN];
	// M^-1 I^T of an element with det J = 1, [m][q]: the volume gradient's projection of the cubature values
	T MIt[kTriMaxNp * kTriMaxNc];
	int p, Np, Nc, N;
};
bool build_tri_ops(int p, TriOps<double>& o);
void tri_geometry(const TriOps<double>& o, const double* xyz, double* met, double* detJ);
void tri_face_nj(int p, int lf, const double* xyz, const double* s, int nq, double* nj);
bool tri_mass_data(const TriOps<double>& o, const double* detJ, double* minv, double* K);

// A face with a TRI element on at least one side, evaluated by k_mix_faces: each side's values at its own face
// points (TRI: its trace at its GL face points, written per element beforehand; QUAD: its GLL face nodes),
// interpolated to the face's GL points, the numerical flux there, and each side's lifted flux
// z = P^T w (J_f F) / w_own at its own GL points, own ordering.  The faces are sorted by (Nf, kL, NL, kR, NR).
struct MixFace {
	long long uL, uR;    // first state value of each side's element
	long long zL, zR;    // first lifted-flux value of each side's own face block [v][a] (Navier-Stokes: then y [d][v][a])
	int kL, kR;          // element type of each side: 0 QUAD, 1 TRI
	int NL, NR;          // p + 1 of each side
	int lfL, lfR;        // local faces
	int rev;             // 1: the right face runs the other way (ind_ord 1)
	int bc;              // DGX_BC_* (boundary faces: the right side is unused)
	int eL, eR;          // each side's element index within its (type, p) group
	int tL, tR;          // TRI sides: the TRI group index
};

// a run of mixed faces with the same number of face points (one launch each)
struct MixRun {
	int Nf, start, count;
};

// The TRI elements of one order of a mixed mesh (a contiguous range of the state vector).
struct TriGroup {
	int p = 0, E = 0, e0 = 0;
	long long u_off = 0;             // first state value
	long long z_off = 0;             // first lifted-flux value (E x 3 x zblk x (p + 1), zblk = nv or 3 nv)
	double* met = nullptr;           // E x 4 x Nc: J-scaled cofactors (y_s, -x_s, -y_r, x_r) at the cubature
	double* detJ = nullptr;          // E x Nc
	double* tr = nullptr;            // E x 3 x nv x N: the state's traces at the own GL face points (written each residual)
	// Navier-Stokes (BR2): every element affine (det J constant): 1 / det J per element with the reference
	// M^-1 and K_f; otherwise each element's own
	int affine = 0;
	double* jinv = nullptr;          // E
	double* minv = nullptr;          // E x Np x Np: inverse mass matrices
	double* K = nullptr;             // E x 3 x N x N: lift traces Ifc_f M^-1 tw0_f per local face
	double* gv = nullptr;            // E x 2 x nv x Np: volume part of the weak gradient (written each residual)
	double* gvt = nullptr;           // E x 3 x 2 x nv x N: its traces at the own GL face points
	// FP32 shadows of the geometry and mass data (the FP32 residual; its traces and gradients use the first half
	// of tr, gv and gvt)
	float* met32 = nullptr;
	float* jinv32 = nullptr;
	float* minv32 = nullptr;
	float* K32 = nullptr;
};

// a run of faces between the same two orders (the faces are sorted by pair)
struct HpPair {
	int NL, NR, start, count;
};

// A variable-order mesh: one uniform-order context per order present; elements sorted by order, so each group
// is a contiguous range of the state vector.
struct HpData {
	std::vector<dgx_ctx*> groups;
	std::vector<int> order;          // p of each group
	std::vector<int> e_off;          // first element of each group (groups + 1 entries)
	std::vector<long long> u_off;    // first state value of each group
	std::vector<long long> z_off;    // first lifted-flux value of each group
	bool ready = false;              // every group uploaded
	int E = 0, F = 0, Fnc = 0;
	long long ndof = 0, nz = 0;
	std::vector<HpPair> pairs;       // runs of faces with one (NL, NR)
	HpFace* faces = nullptr;         // Fnc, sorted by (NL, NR)
	double* nJ = nullptr;            // Fnc x kHpMaxN x 2: n J_f at the face GL points (left ordering, left outward)
	double* nJhi = nullptr;          // the same from the higher-order side's own face (mixed form)
	float* nJ32 = nullptr;
	double* z64 = nullptr;           // lifted fluxes, FP64 and FP32 (each group: E_g x 4 x nv x N_g)
	float* z32 = nullptr;
	HpOps<double>* ops64 = nullptr;
	HpOps<float>* ops32 = nullptr;
	// Side streams (default priority) over which the independent launches of one phase are spread: the
	// groups' kernels and the per-pair face kernels write disjoint outputs, so they run concurrently between
	// a fork (each side stream waits for the caller's stream) and a join (the caller's stream waits for each).
	static constexpr int kLanes = 4;
	hipStream_t lane[kLanes] = {};
	hipEvent_t fork_ev = nullptr;
	hipEvent_t join_ev[kLanes] = {};
	// TRI groups and the faces that touch a TRI element (mixed meshes)
	std::vector<TriGroup> tri;
	TriOps<double>* tri_ops64 = nullptr;   // device: kTriMaxP + 1 entries, indexed by p
	TriOps<float>* tri_ops32 = nullptr;
	MixFace* mfaces = nullptr;             // Fmix, sorted by (Nf, kL, NL, kR, NR)
	double* mnJ = nullptr;                 // Fmix x kHpMaxN x 2: n J_f at the face GL points (left ordering)
	float* mnJ32 = nullptr;
	double* mxy = nullptr;                 // Navier-Stokes rotating walls: GL point coordinates (Fmix x kHpMaxN x 2)
	int Fmix = 0;
	std::vector<MixRun> mruns;             // runs of mixed faces with one Nf
	std::vector<int> mrun_lane;            // side stream of each run (longest-first balance of count Nf^2)
	std::vector<int> group_lane;     // side stream of each QUAD group and of each TRI group (one longest-first
	std::vector<int> tri_lane;       // balance of the groups' costs)
	std::vector<int> pair_lane;      // side stream of each face pair (longest-first balance of count max(NL,NR)^2)
};

// Fork / join of the variable-order side streams on `main` (the context's stream, or the capturing stream).
hipError_t hp_fork(HpData& H, hipStream_t main);
hipError_t hp_join(HpData& H, hipStream_t main);

// Per-element face record: f, neighbour element (-1 on a boundary), and packed
// (neighbour local face | side << 4 | ord << 8 | bc << 12).
struct ElemFace {
	int f;
	int en;
	int packed;
	int pad;
};

struct Ctx {
	dgx_config cfg{};
	int d = 0, p = 0, N = 0, Np = 0, Nc = 0, Nfc = 0, nv = 0;
	int E = 0, F = 0;
	hipStream_t stream = nullptr;
	int device = 0;
	bool mesh_ready = false;
	double* xyz = nullptr;          // E*d*Np
	int* faces = nullptr;           // F*6
	ElemFace* elem_faces = nullptr; // E*2d
	int* order = nullptr;           // E: element processing order (Morton order of centroids)
	double* hmin = nullptr;         // E: minimum straight edge length (compute_min_length_measure)
	int* color_elems = nullptr;     // E: elements grouped by colour (no two face neighbours share a colour)
	std::vector<int> color_off;     // n_colors + 1 offsets into color_elems
	dgx_mlp* mlp = nullptr;         // attached learned preconditioner (not owned)
	double* ns_gvt = nullptr;       // Navier-Stokes: face traces of the volume gradient (E x 2d x d x nv x Nfc)
	double* ns_kappa = nullptr;     // Navier-Stokes: BR2 lift trace factor per face node (E x 2d x Nfc)
	double* wall_xy = nullptr;      // Navier-Stokes with rotating walls: face GL point coordinates (F x Nfc x 2)
	void* ns_gvt_dual = nullptr;    // Navier-Stokes linearisation: ns_gvt and ns_kappa in dual numbers (first use)
	void* ns_kappa_dual = nullptr;
	int rhs_grid[4] = { 0, 0, 0, 0 };   // residual launch grids by TUNE_* (0 = one workgroup per element block)
	HpData* hp = nullptr;           // variable-order mesh (then this context holds no uniform mesh)
	const double* hp_z64 = nullptr; // order group of a variable-order mesh: its part of the lifted fluxes
	const float* hp_z32 = nullptr;
	bool has_nonconforming = false; // some element face is marked kBcNonconforming (then hp_z* must be set)
	double* metrics_vc = nullptr;   // E*d*d*Nc  (m[(i + d*j)*Nc + n] = J dxi_j/dx_i)
	float* metrics_vc32 = nullptr;
	double* detJ_vc = nullptr;      // E*Nc
	float* detJ_vc32 = nullptr;
	double* nJ_fc = nullptr;        // F*Nfc*d   (unit normal * face Jacobian, left element)
	float* nJ_fc32 = nullptr;
	Ops1D<double>* ops64 = nullptr;
	Ops1D<float>* ops32 = nullptr;
	Ops1D<double> h_ops64{};
	Ops1D<float> h_ops32{};
	void* noalloc_scratch = nullptr; // device address of the static scratch array on this context's device
	double* u0 = nullptr;           // RK stage storage (E*Np*nv)
	double* rtmp = nullptr;         // residual scratch (E*Np*nv)
	unsigned long long* dmax = nullptr;  // device reduction slot(s)
	int* dstatus = nullptr;
	double* rk_uout = nullptr;
	// implicit path
	double* blocks = nullptr;       // E*(Np*nv)^2 diagonal blocks
	double* lu = nullptr;           // factored blocks
	int* piv = nullptr;
	size_t nblk = 0;
	std::vector<double> h_tmp;
};

// preconditioner interface of the Newton-GMRES driver: z = P^-1 v
struct Precond {
	virtual ~Precond() {}
	virtual hipError_t setup(Ctx& c, const double* u, const double* inv_dt) = 0;
	virtual hipError_t apply(Ctx& c, const double* v, double* z) = 0;
};
Precond* make_learned_precond(Ctx& c);  // nullptr when no model is attached

// host-side helpers implemented in dgx_ops.cpp
void build_ops(int d, int p, Ops1D<double>& o64, Ops1D<float>& o32);
void gauss_legendre(int n, double* x, double* w);
void gauss_lobatto(int n, double* x, double* w);
std::vector<int> morton_order(int d, int E, int Np, const double* xyz);

// launchers (each returns hipError_t)
hipError_t launch_geometry(Ctx& c);
hipError_t launch_rhs64(Ctx& c, const double* u, double* r, int mode, unsigned long long* dmax);
// r += the volume term (volume = true) or the face term alone, FP64 (k_rhs_parts.hip)
hipError_t launch_rhs_part64(Ctx& c, const double* u, double* r, bool volume);
hipError_t launch_rhs32(Ctx& c, const float* u, float* r);
hipError_t launch_rhs_ns(Ctx& c, const double* u, double* r);   // 2-D Navier-Stokes (BR2), FP64
hipError_t launch_rhs_ns32(Ctx& c, const float* u, float* r);    // the same in FP32
hipError_t launch_rhs_ns_mixed(Ctx& c, const double* u, float* r);   // FP64 state, FP32 perturbation form
// its linearisation (dual numbers): jv = J v (or J v - mass_scale[e] M_e v), one column of every diagonal block
hipError_t launch_ns_jvp(Ctx& c, const double* u, const double* v, double* jv, const double* mass_scale);
hipError_t launch_ns_jvp32(Ctx& c, const double* u, const double* v, double* jv, const double* mass_scale);
hipError_t launch_ns_diag_col(Ctx& c, const double* u, double* blocks, int col);
hipError_t launch_ns_offd_col(Ctx& c, const double* u, double* blocks, int col, bool transposed);
hipError_t launch_rhs_mixed(Ctx& c, const double* u, float* r);
// variable-order meshes: the lifted fluxes of the faces between orders (FP64 / FP32 state); Navier-Stokes also
// the projected BR2 face products, between the groups' gradient (phase 1) and residual (phase 2) kernels
hipError_t launch_hp_faces64(Ctx& c, const double* u);
hipError_t launch_hp_faces32(Ctx& c, const float* u);
hipError_t launch_rhs_ns_phase(Ctx& c, const double* u, double* r, int phase);
hipError_t launch_rhs_ns32_phase(Ctx& c, const float* u, float* r, int phase);
hipError_t launch_rhs_ns_mixed_phase(Ctx& c, const double* u, float* r, int phase);
hipError_t launch_hp_faces_mixed(Ctx& c, const double* u);   // FP64 state and fluxes, FP32 lifted results
// mixed TRI / QUAD meshes (k_tri.hip): one TRI group's residual from its lifted face fluxes; the faces with a TRI side
hipError_t launch_tri_rhs64(const TriGroup& g, const TriOps<double>* ops, const double* u, double* r, const double* z,
                            hipStream_t st);
hipError_t launch_mix_faces64(const HpData& H, const Ctx& c, const double* u, hipStream_t st);
// Navier-Stokes on mixed meshes: a TRI group's volume gradient and its traces (phase 1) or its residual (phase 2),
// and the faces with a TRI side (the BR2 products y, then the lifted inviscid + viscous fluxes)
hipError_t launch_tri_ns_phase64(const TriGroup& g, const TriOps<double>* ops, const double* u, double* r,
                                 const double* z, double mu, double Pr, int phase, hipStream_t st);
// Euler: a TRI group's state traces at its own GL face points (before the mixed faces)
hipError_t launch_tri_trace64(const TriGroup& g, const TriOps<double>* ops, const double* u, hipStream_t st);
hipError_t launch_mix_faces_ns64(const HpData& H, const Ctx& c, const double* u, hipStream_t st);
// the FP32 residual on mixed meshes (FP32 state, geometry shadows and operators; VALU kernels)
hipError_t launch_tri_trace32(const TriGroup& g, const TriOps<float>* ops, const float* u, hipStream_t st);
hipError_t launch_tri_rhs32(const TriGroup& g, const TriOps<float>* ops, const float* u, float* r, const float* z,
                            hipStream_t st);
hipError_t launch_tri_ns_phase32(const TriGroup& g, const TriOps<float>* ops, const float* u, float* r,
                                 const float* z, double mu, double Pr, int phase, hipStream_t st);
hipError_t launch_mix_faces32(const HpData& H, const Ctx& c, const float* u, bool ns);
hipError_t launch_minv(Ctx& c, double* r);
hipError_t launch_positivity(Ctx& c, double* u, int* status);
hipError_t launch_maxabs(Ctx& c, const double* r, size_t n, unsigned long long* dmax);
hipError_t launch_convert(const double* src, float* dst, size_t n, hipStream_t s);
hipError_t launch_convert_back(const float* src, double* dst, size_t n, hipStream_t s);
hipError_t launch_rk_stage(Ctx& c, double* u, const double* r, double dt, int stage);
// the same with the stage-0 copy in u0 (E*Np*nv doubles) instead of c.u0
hipError_t launch_rk_stage_buf(Ctx& c, double* u, double* u0, const double* r, double dt, int stage);
// the same over n values on stream st (a TRI group's range of a mixed mesh)
hipError_t launch_rk_stage_n(hipStream_t st, double* u, double* u0, const double* r, double dt, int stage, size_t n);
// r <- M^-1 r on one TRI group (and max |M^-1 r| into dmax when dmax is not NULL)
hipError_t launch_tri_minv64(const TriGroup& g, const TriOps<double>* ops, double* r, unsigned long long* dmax,
                             hipStream_t st);
// dgx_noalloc.hip: the device address of the static scratch array on device `dev` (the current device);
// looked up once per device, then cached
hipError_t noalloc_scratch(int dev, void** ptr);
hipError_t launch_sub_mass(Ctx& c, double* blocks, const double* scale_e, double scale);
hipError_t launch_mass_sub(Ctx& c, const double* x, const double* scale, double* y);
hipError_t launch_cfl_dt(Ctx& c, const double* u, double cfl, double* inv_dt);
hipError_t launch_relaxed_update(Ctx& c, double* u, const double* du, int* status);
std::vector<double> min_edge_lengths(int d, int N, int E, const double* xyz);
constexpr int kMaxLuBlock = 143;   // largest block of the in-LDS batched LU: nb^2 doubles in the 160 KB of LDS a
                                   // gfx950 workgroup may hold (HEX p = 2: nb = 135)
constexpr int kMaxInvBlock = 142;  // largest element block (nb = Np * nv) of the in-LDS batched inverse: nb^2 + nb
                                   // doubles and the pivot list within the 160 KB of a gfx950 workgroup
constexpr int kMaxSgsFace = 288;   // face dofs (2 d nv N^(d-1)) of one element in the block SGS sweep (HEX p = 2: 270)
hipError_t launch_block_inverse(hipStream_t s, double* blocks, int E, int nb, int* status);
hipError_t launch_block_inv_apply(hipStream_t s, const double* invT, const double* v, double* z, int E, int nb);
hipError_t launch_block_lu(hipStream_t s, double* blocks, int* piv, int E, int nb);
hipError_t launch_block_lu_solve(hipStream_t s, const double* lu, const int* piv, double* x, int E, int nb);
hipError_t launch_diag_col(Ctx& c, const double* u, double* blocks, int col);
// column col < nv * N^(d-1) of the compact face blocks: blocks[((e * 2d + lf) * nfv + col) * nfv + row], nfv =
// nv * N^(d-1), rows and columns on the face nodes (see MODE_OFFD)
hipError_t launch_offd_col(Ctx& c, const double* u, double* blocks, int col);
hipError_t launch_expand_face_blocks(Ctx& c, const double* cT, double* full);   // compact -> full nb x nb (2-D)
// block-sparse assembly: values[slot[e]] (+)= src[e * stride + off] (slot < 0: skipped)
hipError_t launch_bsr_place(hipStream_t s, const double* src, size_t stride, size_t off, const int* slot, double* vals,
                            int E, int nb2, bool add);
// jv = J v, or J v - mass_scale[e] M_e v when mass_scale is not NULL (the pseudo-time Newton operator)
hipError_t launch_jvp(Ctx& c, const double* u, const double* v, double* jv, const double* mass_scale = nullptr);
// the same in FP32 dual numbers (FP32 geometry and operators; FP64 u, v and jv)
hipError_t launch_jvp32(Ctx& c, const double* u, const double* v, double* jv, const double* mass_scale = nullptr);
hipError_t launch_fill_u64(hipStream_t s, unsigned long long* p, int n, unsigned long long v);
hipError_t launch_fill_i32(hipStream_t s, int* p, int n, int v);

// residual modes
enum { MODE_RHS = 0, MODE_RHS_MINV = 1 };
// residual kernels with a tunable launch grid (the first three match DGX_GRAPH_RHS_*)
enum { TUNE_RHS64 = 0, TUNE_RHS32 = 1, TUNE_MIXED = 2, TUNE_RHS64_MINV = 3 };

}  // namespace dgx

struct dgx_ctx {
	dgx::Ctx c;
};

#define DGX_HIP_CHECK(expr)                                  \
	do {                                                     \
		hipError_t e_ = (expr);                              \
		if (e_ != hipSuccess) return e_;                     \
	} while (0)
