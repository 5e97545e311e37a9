// k_dual.hip -- forward-mode (dual number) instantiations of the fused residual kernel:
// exact Jacobian-vector products (FP64, and FP32 for the inner operator of the implicit solve), columns of the
// element diagonal blocks and of the face blocks (Navier-Stokes: J v and the diagonal blocks from k_ns.hip)
// (the quantities constructor_lhs_v_1_T / constructor_lhs_f_1_T assemble,
//  compute_volume_rlhs_T.c:179-220, compute_face_rlhs_T.c:163-220).
#include "k_residual.h"

namespace dgx {

template <int MODE, typename S = Dual, typename R = double>
static hipError_t dispatch_dual(Ctx& c, const RhsArgs<S, R>& a)
{
	if (c.has_nonconforming) return hipErrorInvalidValue;   // no linearisation of faces between orders
	if (c.cfg.pde == DGX_PDE_NAVIER_STOKES) return hipErrorInvalidValue;   // its own kernels (k_ns.hip)
	const int d = c.d, p = c.p, pde = c.cfg.pde, flux = c.cfg.num_flux;
#define DGX_DCASE(NN)                                                                                    \
	if (d == 2 && p + 1 == NN) {                                                                         \
		if (pde == DGX_PDE_ADVECTION) return run_rhs<S, R, 2, NN, DGX_PDE_ADVECTION, DGX_FLUX_UPWIND, MODE>(a, c.stream); \
		if (flux == DGX_FLUX_LF) return run_rhs<S, R, 2, NN, DGX_PDE_EULER, DGX_FLUX_LF, MODE>(a, c.stream);            \
		return run_rhs<S, R, 2, NN, DGX_PDE_EULER, DGX_FLUX_ROE, MODE>(a, c.stream);                                    \
	}
	DGX_DCASE(2) DGX_DCASE(3) DGX_DCASE(4) DGX_DCASE(5)
#undef DGX_DCASE
	// 3-D (HEX): J v, the diagonal blocks and the face blocks, FP64 dual numbers
#define DGX_DCASE3(NN)                                                                                   \
	if (d == 3 && p + 1 == NN) {                                                                         \
		if (pde == DGX_PDE_ADVECTION) return run_rhs<S, R, 3, NN, DGX_PDE_ADVECTION, DGX_FLUX_UPWIND, MODE>(a, c.stream); \
		if (flux == DGX_FLUX_LF) return run_rhs<S, R, 3, NN, DGX_PDE_EULER, DGX_FLUX_LF, MODE>(a, c.stream);            \
		return run_rhs<S, R, 3, NN, DGX_PDE_EULER, DGX_FLUX_ROE, MODE>(a, c.stream);                                    \
	}
	if constexpr (sizeof(S) == 2 * sizeof(double)) { DGX_DCASE3(2) DGX_DCASE3(3) DGX_DCASE3(4) DGX_DCASE3(5) }
#undef DGX_DCASE3
	return hipErrorInvalidValue;
}

static RhsArgs<Dual, double> make_dual(Ctx& c, const double* u, const double* v, double* out, int col)
{
	RhsArgs<Dual, double> a{};
	a.u = u;
	a.v = v;
	a.out = out;
	a.met = c.metrics_vc;
	a.detJ = c.detJ_vc;
	a.nJ = c.nJ_fc;
	a.ef = c.elem_faces;
	a.order = c.order;
	a.ops = c.ops64;
	for (int k = 0; k < 5; ++k) a.bc_prim[k] = c.cfg.bc_prim[k];
	for (int k = 0; k < 3; ++k) a.b_adv[k] = c.cfg.b_adv[k];
	a.E = c.E;
	a.col = col;
	a.dmax = nullptr;
	return a;
}

hipError_t launch_jvp(Ctx& c, const double* u, const double* v, double* jv, const double* mass_scale)
{
	if (c.cfg.pde == DGX_PDE_NAVIER_STOKES) return launch_ns_jvp(c, u, v, jv, mass_scale);
	RhsArgs<Dual, double> a = make_dual(c, u, v, jv, 0);
	a.mass_scale = mass_scale;
	return dispatch_dual<MODE_JVP>(c, a);
}

// the same product in FP32 dual numbers over the FP32 geometry and operator shadows, FP64 vectors in and out
hipError_t launch_jvp32(Ctx& c, const double* u, const double* v, double* jv, const double* mass_scale)
{
	if (c.cfg.pde == DGX_PDE_NAVIER_STOKES) return launch_ns_jvp32(c, u, v, jv, mass_scale);
	RhsArgs<Dual32, float> a{};
	a.u = u;
	a.v = v;
	a.out = jv;
	a.met = c.metrics_vc32;
	a.detJ = c.detJ_vc32;
	a.nJ = c.nJ_fc32;
	a.ef = c.elem_faces;
	a.order = c.order;
	a.ops = c.ops32;
	for (int k = 0; k < 5; ++k) a.bc_prim[k] = c.cfg.bc_prim[k];
	for (int k = 0; k < 3; ++k) a.b_adv[k] = c.cfg.b_adv[k];
	a.E = c.E;
	a.mass_scale = mass_scale;
	return dispatch_dual<MODE_JVP, Dual32, float>(c, a);
}

hipError_t launch_diag_col(Ctx& c, const double* u, double* blocks, int col)
{
	if (c.cfg.pde == DGX_PDE_NAVIER_STOKES) return launch_ns_diag_col(c, u, blocks, col);
	return dispatch_dual<MODE_DIAG>(c, make_dual(c, u, nullptr, blocks, col));
}

hipError_t launch_offd_col(Ctx& c, const double* u, double* blocks, int col)
{
	return dispatch_dual<MODE_OFFD>(c, make_dual(c, u, nullptr, blocks, col));
}

}  // namespace dgx
