// k_rhs_persistent.hip -- persistent-grid instantiations of the fused residual kernel (k_rhs_persistent),
// kept in a code object of their own: next to k_rhs in one translation unit they change k_rhs's register
// allocation.  Launched when the context holds a persistent grid for the kernel (dgx_set_rhs_grid,
// dgx_autotune_rhs).
#include "k_residual.h"

namespace dgx {

hipError_t dispatch_rhs_persistent(const Ctx& c, const RhsArgs<double, double>& a, int mode)
{
	if (mode == MODE_RHS) return dispatch_rhs<double, double, MODE_RHS, true>(c.d, c.p, c.cfg.pde, c.cfg.num_flux, a, c.stream);
	return dispatch_rhs<double, double, MODE_RHS_MINV, true>(c.d, c.p, c.cfg.pde, c.cfg.num_flux, a, c.stream);
}

hipError_t dispatch_rhs_persistent(const Ctx& c, const RhsArgs<float, float>& a, int mode)
{
	if (mode == MODE_PERT) return dispatch_rhs<float, float, MODE_PERT, true>(c.d, c.p, c.cfg.pde, c.cfg.num_flux, a, c.stream);
	return dispatch_rhs<float, float, MODE_RHS, true>(c.d, c.p, c.cfg.pde, c.cfg.num_flux, a, c.stream);
}

}  // namespace dgx
