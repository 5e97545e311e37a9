// k_rhs64.hip -- FP64 instantiations of the fused residual kernel (outer residual, explicit path).
#include "k_residual.h"

namespace dgx {

static RhsArgs<double, double> make_args64(Ctx& c, const double* u, double* r, unsigned long long* dmax)
{
	RhsArgs<double, double> a{};
	a.u = u;
	a.v = nullptr;
	a.out = r;
	a.met = c.metrics_vc;
	a.detJ = c.detJ_vc;
	a.nJ = c.nJ_fc;
	a.ef = c.elem_faces;
	a.order = c.order;
	a.ops = c.ops64;
	for (int k = 0; k < 5; ++k) a.bc_prim[k] = c.cfg.bc_prim[k];
	for (int k = 0; k < 3; ++k) a.b_adv[k] = c.cfg.b_adv[k];
	a.E = c.E;
	a.col = 0;
	a.dmax = dmax;
	a.zf = c.hp_z64;
	return a;
}

hipError_t launch_rhs64(Ctx& c, const double* u, double* r, int mode, unsigned long long* dmax)
{
	if (c.cfg.pde == DGX_PDE_NAVIER_STOKES) {
		// gradient and residual kernels, then M^-1 and the max-norm for the explicit path
		DGX_HIP_CHECK(launch_rhs_ns(c, u, r));
		if (mode != MODE_RHS_MINV) return hipSuccess;
		DGX_HIP_CHECK(launch_minv(c, r));
		return dmax ? launch_maxabs(c, r, (size_t)c.E * c.Np * c.nv, dmax) : hipSuccess;
	}
	RhsArgs<double, double> a = make_args64(c, u, r, dmax);
	if (c.has_nonconforming && !a.zf) return hipErrorInvalidValue;
	a.grid = c.rhs_grid[mode == MODE_RHS ? TUNE_RHS64 : TUNE_RHS64_MINV];
	if (a.grid > 0) return dispatch_rhs_persistent(c, a, mode);
	if (mode == MODE_RHS) return dispatch_rhs<double, double, MODE_RHS>(c.d, c.p, c.cfg.pde, c.cfg.num_flux, a, c.stream);
	return dispatch_rhs<double, double, MODE_RHS_MINV>(c.d, c.p, c.cfg.pde, c.cfg.num_flux, a, c.stream);
}

}  // namespace dgx
