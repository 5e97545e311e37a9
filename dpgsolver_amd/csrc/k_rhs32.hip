// k_rhs32.hip -- FP32 instantiations of the fused residual kernel (defect-correction inner residual):
// FP32 state, FP32 geometry shadow, FP32 arithmetic; and the mixed variant (FP64 state, FP64 pointwise
// fluxes, FP32 contractions in per-element perturbation form).
#include "k_residual.h"

namespace dgx {

hipError_t launch_rhs32(Ctx& c, const float* u, float* r)
{
	RhsArgs<float, float> a{};
	a.u = u;
	a.v = nullptr;
	a.out = r;
	a.met = c.metrics_vc32;
	a.detJ = c.detJ_vc32;
	a.nJ = c.nJ_fc32;
	a.ef = c.elem_faces;
	a.order = c.order;
	a.ops = c.ops32;
	for (int k = 0; k < 5; ++k) a.bc_prim[k] = c.cfg.bc_prim[k];
	for (int k = 0; k < 3; ++k) a.b_adv[k] = c.cfg.b_adv[k];
	a.E = c.E;
	a.dmax = nullptr;
	a.zf = c.hp_z32;
	if (c.has_nonconforming && !a.zf) return hipErrorInvalidValue;
	a.grid = c.rhs_grid[TUNE_RHS32];
	if (a.grid > 0) return dispatch_rhs_persistent(c, a, MODE_RHS);
	return dispatch_rhs<float, float, MODE_RHS>(c.d, c.p, c.cfg.pde, c.cfg.num_flux, a, c.stream);
}

hipError_t launch_rhs_mixed(Ctx& c, const double* u, float* r)
{
	if (c.cfg.pde != DGX_PDE_EULER) return hipErrorInvalidValue;
	RhsArgs<float, float> a{};
	a.u = nullptr;
	a.u64 = u;
	a.v = nullptr;
	a.out = r;
	a.met = c.metrics_vc32;
	a.detJ = c.detJ_vc32;
	a.nJ = c.nJ_fc32;
	a.ef = c.elem_faces;
	a.order = c.order;
	a.ops = c.ops32;
	for (int k = 0; k < 5; ++k) a.bc_prim[k] = c.cfg.bc_prim[k];
	for (int k = 0; k < 3; ++k) a.b_adv[k] = c.cfg.b_adv[k];
	a.E = c.E;
	a.dmax = nullptr;
	a.zf = c.hp_z32;
	if (c.has_nonconforming && !a.zf) return hipErrorInvalidValue;
	a.grid = c.rhs_grid[TUNE_MIXED];
	if (a.grid > 0) return dispatch_rhs_persistent(c, a, MODE_PERT);
	return dispatch_rhs<float, float, MODE_PERT>(c.d, c.p, c.cfg.pde, c.cfg.num_flux, a, c.stream);
}

}  // namespace dgx
