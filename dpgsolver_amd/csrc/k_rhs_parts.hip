// k_rhs_parts.hip -- FP64 volume-only and face-only instantiations of the fused residual kernel
// (compute_volume_rlhs_dg / compute_face_rlhs_dg called separately, as the reference's linearization harness
// does, src/testing/integration/test_integration_linearization.c:270-271).  A translation unit of their own:
// the default residual kernels keep the register allocation they have in k_rhs64.hip.
#include "k_residual.h"

namespace dgx {

hipError_t launch_rhs_part64(Ctx& c, const double* u, double* r, bool volume)
{
	RhsArgs<double, double> a{};
	a.u = u;
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
	a.zf = c.hp_z64;
	if (c.has_nonconforming) return hipErrorInvalidValue;   // order groups of a variable-order mesh: not served
	if (volume) return dispatch_rhs<double, double, MODE_VOL>(c.d, c.p, c.cfg.pde, c.cfg.num_flux, a, c.stream);
	return dispatch_rhs<double, double, MODE_FACE>(c.d, c.p, c.cfg.pde, c.cfg.num_flux, a, c.stream);
}

}  // namespace dgx
