/*
 * dg_oracle.h -- CPU FP64 restatement of DPGSolver's DG residual path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path (libdgx).  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it.  The product never links or calls it.
 *
 * What it restates (reference paths relative to the DPGSolver source tree):
 *   - Gauss-Legendre / Gauss-Lobatto nodes and weights (src/element/nodes.c:69-104;
 *     pinned by input/testing/unit/nodes/nodes_tp.data)
 *   - r-fastest Kronecker assembly of tensor-product operators
 *     (src/element/element_operators_tp.c:90-131; pinned by
 *     input/testing/unit/operators/operators_tp.data)
 *   - metric terms: 2-D cofactors (src/simulation/geometry/geometry_T.c:546-590),
 *     3-D symmetric curl form (deprecated/src/setup_geom_factors.c:157-243),
 *     metrics at GLL "vm" nodes interpolated to volume/face cubature
 *     (geometry_T.c:734-801), face unit normals and face Jacobian
 *     (geometry_T.c:412-445)
 *   - Euler / advection physical flux (src/simulation/test_case/flux/flux_euler_T.c:71-143,
 *     flux_advection_T.c:63)
 *   - Lax-Friedrichs, Roe-Pike, upwind numerical fluxes
 *     (src/simulation/solvers/numerical_flux/numerical_flux_euler_T.c:60-327, 470-778;
 *      numerical_flux_advection_T.c:39-124)
 *   - Riemann-invariant and slip-wall boundary states
 *     (src/simulation/test_case/boundary/boundary_euler_T.c:99-180, 494-530)
 *   - volume term  rhs += sum_d tw1_vt_vc[d] * fr_d   (compute_volume_rlhs_T.c:145-177)
 *   - face term    rhs_L -= tw0_vt_fc * (nnf*J_f); rhs_R -= tw0_vt_fc * (-perm(nnf*J_f))
 *     (compute_face_rlhs_T.c:297-359, compute_face_rlhs.c:66-118)
 *   - dense mass matrix and its inverse (src/simulation/solvers/volume_solver_T.c:115-145),
 *     scale_rhs_by_m_inv_std (solvers/dg/solve_dg.c:265-272)
 *   - SSP-RK3 stage arithmetic (solvers/solve_explicit.c:252-301) and the
 *     Bezier positivity limiter (solvers/solve.c:187-233, 356-374)
 *   - max |rhs|_inf (solvers/compute_rlhs.c:59-71)
 *   - block Jacobian by complex step, the reference's own linearization check
 *     (testing/support/simulation/solvers/dg/test_support_solve_dg.c:83-212)
 *
 * Every routine exists twice: Type=double (suffix _d) and Type=double complex
 * (suffix _c), mirroring the reference's real/complex template twin.
 */
#ifndef DG_ORACLE_H
#define DG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OR_EPS      1.0e-15   /* definitions_tol.h:23 */
#define OR_EPS_PHYS 1.0e-13   /* definitions_physics.h:30 */
#define OR_GAMMA    1.4       /* definitions_physics.h:23 */
#define OR_GM1      0.4
#define OR_GM3     -1.6

enum { OR_PDE_ADVECTION = 0, OR_PDE_EULER = 1 };
enum { OR_FLUX_UPWIND = 0, OR_FLUX_LF = 1, OR_FLUX_ROE = 2 };
enum { OR_BC_INTERIOR = 0, OR_BC_RIEMANN = 1, OR_BC_SLIPWALL = 2,
       /* Navier-Stokes no-slip walls (definitions_bc.h: BC_NOSLIP_ADIABATIC 111, BC_NOSLIP_DIABATIC 112,
        * BC_NOSLIP_ALL_ROTATING 113) */
       OR_BC_NOSLIP_ADIABATIC = 3, OR_BC_NOSLIP_DIABATIC = 4, OR_BC_NOSLIP_ROTATING = 5 };

typedef struct or_ctx or_ctx;

/* 1-D building blocks */
int  or_gauss_legendre(int n, double* x, double* w);
int  or_gauss_lobatto(int n, double* x, double* w);
/* Iop[r*nin+c] = l_c(xout[r]), Dop[r*nin+c] = l_c'(xout[r]) (row-major nout x nin) */
void or_lagrange(int nin, const double* xin, int nout, const double* xout, double* Iop, double* Dop);
/* out = ops[d-1] (x) ... (x) ops[0]; r-fastest index ordering, row-major output */
void or_kron(int d, const double* const* ops, const int* rows, const int* cols, double* out);
/* Bernstein basis on [-1,1] evaluated at x: B[r*(p+1)+j] = B_j(x_r) */
void or_bernstein(int p, int n, const double* x, double* B);

/*
 * Context.  faces: F x 6 ints {eL, fL, eR, fR, ord, bc}; xyz: E x d x Np geometry
 * node coordinates at the GLL solution nodes (r-fastest).
 */
or_ctx* or_create(int d, int p, int pde, int flux, int E, int F,
                  const double* xyz, const int* faces,
                  const double* b_adv, const double* bc_prim);
void or_destroy(or_ctx* c);
int  or_np(const or_ctx* c);
int  or_nv(const or_ctx* c);
void or_geometry(const or_ctx* c, double* metrics_vc, double* detJ_vc, double* nJ_fc);
void or_face_perm(const or_ctx* c, int ord, int* perm);

/* residual (volume + face, no M^-1) -- compute_rlhs_common_dg */
void or_compute_rhs_d(const or_ctx* c, const double* u, double* rhs);
/* Navier-Stokes with BR2 (real and complex; interior, Riemann and no-slip wall faces).  or_set_viscous switches an Euler
 * context to Navier-Stokes with constant viscosity mu and Prandtl number Pr. */
void or_set_viscous(or_ctx* c, double mu, double Pr);
/* wall data of the no-slip conditions: nf_E the normal viscous energy flux of the adiabatic (0) and diabatic
 * walls (get_normal_flux_Energy: -mu Cp / Pr dT/dn); omega, rho_b, E_b of the rotating wall (solution.data) */
void or_set_walls(or_ctx* c, double nf_E, double omega, double rho_b, double E_b);
/* viscosity by Sutherland's formula mu_s = 8e-2 mu sqrt(T) / (1 + 112 / T), T = p / (rho r_s), instead of the
 * constant mu (compute_mu_sutherland_T, solution_navier_stokes_T.c:246-255) */
void or_set_sutherland(or_ctx* c, double r_s);
/* negated viscous flux fv[dim + d*eq] for state s[nv] and gradient g[v + nv*dim] */
void or_viscous_flux(int d, const double* s, const double* g, double mu, double Pr, double* fv);
/* BR2 weak gradient coefficients: g, gv: E x d x nv x Np; gf: F x 2 x d x nv x Np.  0 or -1 */
int or_ns_gradient(or_ctx* c, const double* u, double* g, double* gv, double* gf);
/* full Navier-Stokes residual (no M^-1).  0 or -1 */
int or_compute_rhs_ns(or_ctx* c, const double* u, double* rhs);
/* its complex twin: R(ure + i uim) = rre + i rim (complex-step derivatives of the BR2 residual).  0 or -1 */
int or_compute_rhs_ns_c(or_ctx* c, const double* ure, const double* uim, double* rre, double* rim);
/* the same on nthreads OpenMP threads (faces lifted by colour); agrees with or_compute_rhs_d to rounding */
void or_compute_rhs_d_threads(const or_ctx* c, const double* u, double* rhs, int nthreads);
/* complex twin: u = ure + i*uim */
void or_compute_rhs_c(const or_ctx* c, const double* ure, const double* uim, double* rre, double* rim);
void or_compute_rhs_c_threads(const or_ctx* c, const double* ure, const double* uim, double* rre, double* rim,
                              int nthreads);
/* pieces of the CPU Newton-GMRES (solve_implicit.c): element diagonal blocks by complex step on nthreads threads,
 * the per-element 1/dt of the CFL-constrained pseudo-time step, the relaxed update, dense block mat-vecs */
void or_diag_blocks_cs_threads(const or_ctx* c, const double* u, double h, double* blocks, int nthreads);
void or_cfl_inv_dt(const or_ctx* c, const double* u, double cfl, double* inv_dt);
int or_relaxed_update(const or_ctx* c, double* u, const double* du);
void or_block_apply(int E, int nb, const double* A, const double* v, double* z, int nthreads);
int or_block_inverse(int E, int nb, const double* A, double* Ainv, int nthreads);
void or_dots(int k, long long n, const double* V, const double* w, double* h, int nthreads);
void or_gemv_sub(int k, long long n, const double* V, const double* h, double* w, int nthreads);
/* volume-only and face-only parts */
void or_volume_rhs_d(const or_ctx* c, const double* u, double* rhs);
void or_face_rhs_d(const or_ctx* c, const double* u, double* rhs);

/* rhs <- M^-1 rhs, dense inverse per element */
int  or_apply_minv(or_ctx* c, double* rhs);
/* mass matrix of element e (Np x Np row-major) */
void or_mass(const or_ctx* c, int e, double* M);
/* compute_rhs_dg: residual, M^-1, returns max |rhs|_inf */
double or_compute_rhs_explicit(or_ctx* c, const double* u, double* rhs);
/* positivity limiter applied in place to every element; returns 0 or -1 (mean <= EPS_PHYS) */
int  or_enforce_positivity(const or_ctx* c, double* u);
/* one SSP-RK3 step; returns max |rhs| of the last stage; status via *err */
double or_ssp_rk3_step(or_ctx* c, double* u, double dt, int positivity, int* err);

/* scalar kernels exposed for unit tests (Type=double twin) */
void or_euler_flux_d(int d, const double* s, double* f);                 /* f[dim + d*eq] */
void or_numflux_d(int d, int pde, int flux, const double* sL, const double* sR,
                  const double* n, const double* b_adv, double* nnf);
void or_numflux_c(int d, int pde, int flux, const double* sLre, const double* sLim,
                  const double* sRre, const double* sRim, const double* n,
                  const double* b_adv, double* nre, double* nim);
void or_bc_state_d(int d, int bc, const double* sL, const double* n, const double* bc_prim, double* sR);

/*
 * Jacobian by complex step, step h (reference uses 1e-30).  Returns the
 * (Np*nv) x (Np*nv) block d rhs_{e_row} / d u_{e_col}, row-major, row index
 * eq*Np+node, column index var*Np+node (ind_dof convention, solve_T.c:184-223).
 */
void or_jacobian_block_cs(const or_ctx* c, const double* u, int e_row, int e_col, double h, double* blk);
/* all diagonal blocks (E x (Np nv)^2) */
void or_diag_blocks_cs(const or_ctx* c, const double* u, double h, double* blocks);

#ifdef __cplusplus
}
#endif
#endif
