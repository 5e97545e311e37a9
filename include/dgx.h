/*
 * dgx.h -- C ABI of the MI355X-native DG residual library (libdgx.so).
 *
 * Drop-in boundary for DPGSolver's DG hot path.  The reference entry points each
 * function replaces are cited per declaration (paths relative to the DPGSolver
 * source tree).  All device pointers are HIP device memory on the context's device;
 * every call is asynchronous on the context's stream unless stated otherwise and
 * returns a dgx_status (never aborts, unlike the reference's EXIT_ERROR/assert).
 *
 * State layout at the ABI is the reference's: per element the (Np, nv) column-major
 * block sol_coef (index var*Np + node, nodes r-fastest), elements concatenated in
 * ind_dof order (src/simulation/solvers/solve_T.c:184-223).
 */
#ifndef DGX_H
#define DGX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int dgx_status;
enum {
	DGX_OK = 0,
	DGX_ERR_ARG = -1,          /* bad argument / size */
	DGX_ERR_HIP = -2,          /* HIP runtime error */
	DGX_ERR_UNSUPPORTED = -3,  /* (d, p, pde) not compiled */
	DGX_ERR_POSITIVITY = -4,   /* mean density or pressure <= EPS_PHYS (solve.c:204-208) */
	DGX_ERR_NONFINITE = -5,    /* NaN/Inf in a residual norm */
	DGX_ERR_NOT_CONVERGED = -6,/* linear solve hit its iteration cap (KSP DIVERGED_ITS) */
	DGX_ERR_NO_MESH = -7
};

/* DGX_PDE_NAVIER_STOKES: Euler + BR2 viscous terms (2-D QUAD; FP64, FP32 and mixed residual, FP64 explicit
 * path), constant viscosity mu and Prandtl number Pr from the config.  Boundary faces: Riemann and the no-slip
 * walls, on uniform-order, variable-order and mixed TRI/QUAD meshes. */
enum { DGX_PDE_ADVECTION = 0, DGX_PDE_EULER = 1, DGX_PDE_NAVIER_STOKES = 2 };
/* num_flux_1st: upwind | Lax-Friedrichs | Roe-Pike (numerical_flux/definitions_numerical_flux.h:22-32) */
enum { DGX_FLUX_UPWIND = 0, DGX_FLUX_LF = 1, DGX_FLUX_ROE = 2 };
/* Boundary conditions (definitions_bc.h: BC_RIEMANN 1, BC_SLIPWALL 2; the Navier-Stokes no-slip walls
 * BC_NOSLIP_ADIABATIC 111, BC_NOSLIP_DIABATIC 112, BC_NOSLIP_ALL_ROTATING 113, boundary_navier_stokes_T.c).  The
 * slip wall is an Euler condition, the no-slip walls Navier-Stokes ones (face_solver_T.c:251-298): another
 * combination is refused with DGX_ERR_UNSUPPORTED. */
enum { DGX_BC_INTERIOR = 0, DGX_BC_RIEMANN = 1, DGX_BC_SLIPWALL = 2,
       DGX_BC_NOSLIP_ADIABATIC = 3, DGX_BC_NOSLIP_DIABATIC = 4, DGX_BC_NOSLIP_ROTATING = 5 };

typedef struct dgx_config {
	int d;              /* 2 or 3 (DIM, a compile-time constant in the reference) */
	int p;              /* polynomial order, 1..7 (p_ref) */
	int pde;            /* DGX_PDE_* (pde_name) */
	int num_flux;       /* DGX_FLUX_* (num_flux_1st) */
	double b_adv[3];    /* advection velocity (solution.data b_adv) */
	double bc_prim[5];  /* free-stream primitive state for Riemann BCs (rho, u, v, [w], p) */
	double mu;          /* DGX_PDE_NAVIER_STOKES: constant viscosity (compute_mu_constant_T, solution.data mu) */
	double Pr;          /* DGX_PDE_NAVIER_STOKES: Prandtl number (solution.data Pr) */
	/* no-slip wall data (solution.data; boundary_navier_stokes_T.c): */
	double wall_nf_E;   /* DGX_BC_NOSLIP_ADIABATIC / _DIABATIC: the normal viscous energy flux, which replaces the
	                       energy equation's numerical flux there (get_normal_flux_Energy: 0 adiabatic,
	                       -mu Cp / Pr dTdn diabatic); pass 0 for adiabatic walls */
	double wall_omega;  /* DGX_BC_NOSLIP_ROTATING: angular velocity about the origin (omega); the wall velocity at
	                       (x, y) is omega (-y, x) */
	double wall_rho;    /* DGX_BC_NOSLIP_ROTATING: wall density and total energy (rho_b, E_b) */
	double wall_E;
	int viscosity_type; /* DGX_VISC_CONSTANT (0: mu) or DGX_VISC_SUTHERLAND (solution.data viscosity_type) */
	double r_s;         /* DGX_VISC_SUTHERLAND: specific gas constant of T = p / (rho r_s) (solution.data r_s;
	                       0 means 1) */
} dgx_config;
/* Viscosity of the Navier-Stokes flux: constant mu (compute_mu_constant_T), or Sutherland's formula
 * mu = 8e-2 mu sqrt(T) / (1 + 112 / T) (compute_mu_sutherland_T, solution_navier_stokes_T.c:246-255; uniform-order
 * QUAD meshes, FP64 / FP32 / mixed residuals and the linearisation; variable-order and mixed TRI/QUAD meshes
 * refuse it with DGX_ERR_UNSUPPORTED). */
enum { DGX_VISC_CONSTANT = 0, DGX_VISC_SUTHERLAND = 1 };

typedef struct dgx_mesh {
	int E;              /* elements (QUAD/HEX, one p) */
	int F;              /* faces */
	const double* xyz;  /* HOST, E x d x Np: isoparametric geometry nodes at the GLL solution nodes */
	const int* faces;   /* HOST, F x 6: eL, fL, eR, fR, ord, bc  (face.h:32-70 neigh_info): eL/fL = neigh_info[0]
	                       {volume, ind_lf}, eR/fR = neigh_info[1]; ord = the face orientation index ind_ord, the
	                       case of the reference's face node correspondence (constructor_face_corr,
	                       element/nodes_correspondence.c; input/testing/unit/nodes/nodes_correspondence.data)
	                       that takes the right trace into the left ordering, s_l[i] = s_r[perm[i]]
	                       (permute_Multiarray_T_fc, compute_face_rlhs_T.c:312-318; dgx_face_permutation gives
	                       perm).  2-D: 0 aligned, 1 reversed (Face::neigh_info[0].ind_ord, equal to
	                       neigh_info[1].ind_ord for LINE faces).  3-D: 0..7, Face::neigh_info[0].ind_ord (from
	                       the left-to-right vertex match, face.c:498-540; neigh_info[1] holds the inverse case,
	                       which differs only for cases 5 and 6).
	                       Boundary faces: -1 (as the reference stores them) or any valid code, ignored.  Any
	                       other value: DGX_ERR_ARG. */
} dgx_mesh;

/* 2-D QUAD mesh with one order per element (Solver_Volume_T::p_ref, volume_solver_T.h:31-104). */
typedef struct dgx_mesh_hp {
	int E;              /* elements */
	int F;              /* faces */
	const int* p;       /* HOST, E orders 1 <= p[e] <= dgx_config.p, non-decreasing (elements grouped by order) */
	const double* xyz;  /* HOST, per element 2 x (p[e]+1)^2 GLL geometry nodes, element after element */
	const int* faces;   /* HOST, F x 6 as dgx_mesh.faces */
} dgx_mesh_hp;

/* 2-D mesh of QUAD and TRI elements of one order each (the reference's mixed meshes, mesh_type mixed): element e
 * has type[e] (0 QUAD: Lagrange basis on the (p+1)^2 GLL nodes; 1 TRI: Lagrange basis on the (p+1)(p+2)/2 WSH
 * nodes of the reference's equilateral triangle, interp_si WSH, element/nodes.c:628-852; dgx_reference_nodes
 * gives both node sets) and order p[e]; the elements are sorted by (type, p), so every (type, p) group is one
 * contiguous range of the state.  xyz: per element 2 x Np_e geometry nodes at the solution nodes (isoparametric),
 * element after element.  faces: F x 6 as dgx_mesh.faces; TRI local face lf is opposite vertex lf and runs from
 * vertex (1, 2, 0)[lf] to (2, 2, 1)[lf] (element.c:454-463); ord 1 = the right face runs the other way. */
typedef struct dgx_mesh_mixed {
	int E;
	int F;
	const int* type;    /* HOST, E: 0 QUAD, 1 TRI */
	const int* p;       /* HOST, E: orders 1..5 */
	const double* xyz;  /* HOST, per element 2 x Np_e */
	const int* faces;   /* HOST, F x 6 */
} dgx_mesh_mixed;

typedef struct dgx_ctx dgx_ctx;
typedef struct dgx_mlp dgx_mlp;

/* Host mesh read from a file (dgx_mesh_read_gmsh); arrays are malloc'ed, release with dgx_mesh_free. */
typedef struct dgx_host_mesh {
	int d, p, E, F;
	double* xyz;          /* E x d x (p+1)^d straight-sided geometry at the GLL nodes (dgx_mesh.xyz) */
	int* faces;           /* F x 6 {eL, fL, eR, fR, ord, bc} (dgx_mesh.faces) */
	int* v_to_v;          /* E x 2d: neighbour across each local face, -1 on a boundary (Mesh_Connectivity) */
	int* v_to_lf;         /* E x 2d: the neighbour's local face, or the boundary's physical tag */
	int* v_to_lf_wp;      /* E x 2d: as v_to_lf, with periodic faces holding their periodic boundary tag */
	int n_unsupported_bc; /* boundary faces whose tag is neither Riemann (x101) nor slip wall (x102) */
	int n_inverted;       /* elements numbered clockwise (det J <= 0 in the straight map) */
} dgx_host_mesh;

/* Host only: perm[i] (N^(d-1) ints, N = p + 1) = the right-face node that coincides with left-face node i for
 * orientation code ord of the face table, as the library applies it (constructor_face_corr,
 * element/nodes_correspondence.c).  DGX_ERR_ARG for a code that means nothing in dimension d. */
dgx_status dgx_face_permutation(int d, int p, int ord, int* perm);

/* Library/version and device helpers. */
const char* dgx_version(void);
dgx_status dgx_device_count(int* n);

/* Context for one discretisation.  Replaces constructor_Simulation's solver set-up
 * (src/simulation/simulation.h:214-216) for the DG hot path. */
dgx_status dgx_create(dgx_ctx** ctx, const dgx_config* cfg);
dgx_status dgx_destroy(dgx_ctx* ctx);
/* hipStream_t; NULL = the default stream */
dgx_status dgx_set_stream(dgx_ctx* ctx, void* stream);
dgx_status dgx_synchronize(dgx_ctx* ctx);
dgx_status dgx_sizes(const dgx_ctx* ctx, int* Np, int* nv, int* Nfc, long long* ndof);

/* Mesh upload + FP64 geometry on the GPU: volume metrics (2-D cofactors, 3-D symmetric
 * curl form), det J at the volume cubature, n*J_f at the face cubature, FP32 shadows.
 * Replaces set_up_solver_geometry_T / compute_geometry_volume_T / compute_geometry_face_T
 * (src/simulation/geometry/geometry_T.c:165, 208-325, 734-801;
 *  deprecated/src/setup_geom_factors.c:157-243 for the 3-D curl form). Synchronous. */
dgx_status dgx_upload_mesh(dgx_ctx* ctx, const dgx_mesh* mesh);
/* Variable-order mesh (2-D QUAD, Euler and Navier-Stokes): element e has order p[e] and its state block
 * nv x (p[e]+1)^2 in the layout of dgx_mesh, blocks in element order.  A face takes the larger order of its
 * two volumes (compute_face_reference_order, face_solver.c:123-131) and its geometry from the lower-order side
 * (get_dominant_geom_vol_side_index, face_solver.c:95-107).  The context must be created with
 * dgx_config.p = the largest order allowed.  It then serves dgx_compute_rhs, dgx_compute_rhs_f32,
 * dgx_compute_rhs_mixed, dgx_compute_rhs_explicit, dgx_ssp_rk3_step, dgx_graph_create_rhs (the face kernel
 * between orders and every order group's launches in one graph) and dgx_sizes (ndof = the total); its other entry
 * points return DGX_ERR_UNSUPPORTED.  Navier-Stokes: interior (periodic) faces only.  The context then owns
 * four side streams over which each residual spreads the independent per-order and per-pair launches, forked
 * from and joined back to the context's stream within the call.  Replaces the per-(type,
 * p) volume/face loops of compute_volume_rlhs_dg / compute_face_rlhs_dg for meshes whose volumes differ in
 * p_ref.  Synchronous. */
dgx_status dgx_upload_mesh_hp(dgx_ctx* ctx, const dgx_mesh_hp* mesh);
/* Mixed TRI / QUAD mesh (Euler).  TRI groups: the volume term with the element's dense operators at the WV
 * cubature of degree 2p (compute_volume_rlhs_T.c:145-177; nodes_operators.c:327-345, 436-520); every face with a
 * TRI side is evaluated once at the GL points of the face order max(p_L, p_R) (compute_face_reference_order,
 * face_solver.c:123-131), geometry from the lower-order side, and lifted with each side's own test functions.
 * QUAD groups and faces between QUADs as in dgx_upload_mesh_hp.  Serves dgx_compute_rhs (FP64), dgx_graph_create_rhs
 * (DGX_GRAPH_RHS_FP64) and dgx_sizes; with TRI elements present the FP32, mixed and explicit entry points return
 * DGX_ERR_UNSUPPORTED, as does a Navier-Stokes context.  Elements of negative Jacobian: DGX_ERR_ARG.
 * Synchronous. */
dgx_status dgx_upload_mesh_mixed(dgx_ctx* ctx, const dgx_mesh_mixed* mesh);
/* Host only: the solution nodes of the reference element of type (0 QUAD, 1 TRI) and order p (1..5): rst (2 Np,
 * (r, s) pairs), weights w (Np; the WSH rule's weights on the triangle of area sqrt(3), or the GLL tensor weights),
 * *np = Np; any output may be NULL. */
dgx_status dgx_reference_nodes(int type, int p, double* rst, double* w, int* np);
/* Host only: the reference's symmetric TRI rule kind<q> (kind 0: WSH of order q, 1: WV of degree q) expanded as the
 * reference expands it (nodes.c:628-852): *n points rst (2 n) and weights w (n); NULL outputs are skipped. */
dgx_status dgx_tri_rule(int kind, int q, double* rst, double* w, int* n);
/* gmsh 2.2 ASCII reader for 2-D QUAD and 3-D HEX meshes (mesh/mesh_readers.c, mesh_connectivity.c:285-465,
 * mesh_periodic.c): volumes in file order, periodic pairs from the tagged boundary faces (51/52, 53/54, 55/56
 * master/slave in x, y, z), boundary tags from the line (2-D) or QUAD (3-D) elements, each interior face's
 * ind_ord from its corners; geometry is the straight-sided (bilinear / trilinear) map at the GLL nodes of degree
 * p.  TRI, TET, prism and pyramid volumes: DGX_ERR_UNSUPPORTED.  Host only. */
dgx_status dgx_mesh_read_gmsh(const char* path, int p, dgx_host_mesh* out);
void dgx_mesh_free(dgx_host_mesh* m);
/* A gmsh 2.2 mesh of TRI and QUAD elements (the reference's mixed 2-D meshes, mesh_readers.c,
 * mesh_connectivity.c:285-465): the volumes in file order.  TRI local face f is the edge opposite vertex f,
 * its vertices (1, 2), (0, 2), (0, 1) in increasing face coordinate; a TRI's geometry is the affine map of its
 * vertices at the WSH solution nodes.  v_to_v / v_to_lf / v_to_lf_wp hold 4 entries per volume (a TRI's 4th is
 * -1).  dgx_upload_mesh_mixed wants the volumes sorted by (type, p): reorder them and renumber the face table
 * first.  Host only; release with dgx_mesh_mixed_free. */
typedef struct dgx_host_mesh_mixed {
	int E, F, p;
	int* type;            /* E: 0 QUAD, 1 TRI */
	double* xyz;          /* per volume 2 x Np (x then y at its solution nodes), concatenated */
	int* faces;           /* F x 6 {eL, fL, eR, fR, ord, bc} */
	int* v_to_v;          /* E x 4 */
	int* v_to_lf;         /* E x 4 */
	int* v_to_lf_wp;      /* E x 4 */
	int n_unsupported_bc;
	int n_inverted;
} dgx_host_mesh_mixed;
dgx_status dgx_mesh_read_gmsh_mixed(const char* path, int p, dgx_host_mesh_mixed* out);
void dgx_mesh_mixed_free(dgx_host_mesh_mixed* m);
/* The reference's restart file (restart_writers.c:57-171, 301-412): gmsh 2.2 vertices, boundary faces tagged
 * 10101 (Riemann) / 10102 (slip wall), volumes, and $SolutionCoefficients with each volume's Bezier
 * coefficients (dof fastest).  u: HOST state in the dgx_compute_rhs layout.  Host only. */
dgx_status dgx_restart_write_gmsh(const char* path, const dgx_config* cfg, const dgx_mesh* mesh, const double* u);
/* Reads $SolutionCoefficients back into nodal values (initialize_volumes_sol_coef, restart_T.c:317-420);
 * a different degree in the file: DGX_ERR_UNSUPPORTED. */
dgx_status dgx_restart_read_solution(const char* path, const dgx_config* cfg, int E, double* u);
/* Copies of the device geometry into HOST arrays (any may be NULL). Synchronous. */
dgx_status dgx_get_geometry(dgx_ctx* ctx, double* metrics_vc, double* detJ_vc, double* nJ_fc);

/* FP64 residual r = R(u): volume + face terms, no M^-1.  Replaces compute_rlhs_common_dg
 * (solvers/dg/solve_dg.c:182-189) = compute_volume_rlhs_dg + compute_face_rlhs_dg with
 * 'e' parameters (solvers/dg/compute_volume_rlhs_dg_T.c:53-72, compute_face_rlhs_dg_T.c:88-114).
 * r is fully overwritten (no pre-zeroing needed, solve_T.c:159-182). */
dgx_status dgx_compute_rhs(dgx_ctx* ctx, const double* u, double* r);
/* r += the volume term alone: compute_volume_rlhs_dg with 'e' parameters (solvers/dg/compute_volume_rlhs_dg_T.c:53-72,
 * compute_volume_rlhs_T.c:145-177), rhs += sum_d tw1_vt_vc[d] fr_d.  Like the reference, it accumulates into r:
 * the caller zeroes r first (solve_T.c:159-182).  FP64, uniform-order Euler and advection contexts. */
dgx_status dgx_compute_volume_rhs(dgx_ctx* ctx, const double* u, double* r);
/* r += the face term alone: compute_face_rlhs_dg with 'e' parameters (solvers/dg/compute_face_rlhs_dg_T.c:88-114,
 * compute_face_rlhs_T.c:297-359), rhs -= tw0_vt_fc (J_f nnf) on both sides of every face, boundary faces
 * included.  Accumulates like dgx_compute_volume_rhs; the two calls on a zeroed r give dgx_compute_rhs's r to
 * rounding (the reference's linearization harness calls them in this order,
 * testing/integration/test_integration_linearization.c:270-271). */
dgx_status dgx_compute_face_rhs(dgx_ctx* ctx, const double* u, double* r);
/* FP32 residual (defect-correction inner residual): FP32 state, FP32 geometry shadow,
 * FP32 arithmetic.  Same semantics as dgx_compute_rhs. */
dgx_status dgx_compute_rhs_f32(dgx_ctx* ctx, const float* u, float* r);
/* FP32 residual of an FP64 state (Euler and Navier-Stokes), same semantics as dgx_compute_rhs: the kernel forms
 * u - c_e and F(u) - F(c_e) per element (c_e = the state at the element's first node), evaluates the
 * pointwise physical and numerical fluxes in FP64 and runs the interpolations, contractions and lifts
 * in FP32 on those perturbations.  The F(c_e) part cancels exactly (free-stream preservation), so
 * FP32 rounding acts only on the within-element variation. */
dgx_status dgx_compute_rhs_mixed(dgx_ctx* ctx, const double* u, float* r);
/* compute_rhs_dg (solve_dg.c:130-136): r = M^-1 R(u), and max|r|_inf to *max_abs_r
 * (HOST pointer, may be NULL; synchronous when given) (compute_rlhs.c:59-71). */
dgx_status dgx_compute_rhs_explicit(dgx_ctx* ctx, const double* u, double* r, double* max_abs_r);
/* r <- M^-1 r, M = cv0^T diag(w J) cv0 (scale_rhs_by_m_inv_std, solve_dg.c:265-272). */
dgx_status dgx_apply_minv(dgx_ctx* ctx, double* r);
/* Positivity limiter in place (enforce_positivity_highorder, solve.c:187-233).
 * Returns DGX_ERR_POSITIVITY if an element mean density/pressure is below EPS_PHYS. Synchronous. */
dgx_status dgx_enforce_positivity(dgx_ctx* ctx, double* u);
/* One SSP-RK3 step in place (time_step_ssp_rk_33, solve_explicit.c:252-301), positivity
 * after every stage when enforce_positivity != 0; *max_abs_r (HOST, may be NULL) receives
 * max|M^-1 R| of the last stage.  Synchronous when max_abs_r is given. */
dgx_status dgx_ssp_rk3_step(dgx_ctx* ctx, double* u, double dt, int enforce_positivity, double* max_abs_r);

/* dgx_ssp_rk3_step without any allocation: no counterpart in the reference.  For callers that cannot allocate
 * device memory (code under HIP graph capture, tight time loops, harnesses that forbid hipMalloc): the stage-0
 * copy, the residual, the max-norm slot and the limiter flag live in a static device array of 64 MiB per GPU
 * (held until the process ends) instead of buffers the library allocates on the first dgx_ssp_rk3_step.  Same
 * kernels and results as dgx_ssp_rk3_step.  Uniform-order Euler or advection contexts with at most 4,194,272
 * state values (E * Np * nv: two stage buffers, each rounded up to 256 B, after a 256 B head); larger:
 * DGX_ERR_ARG; variable-order or Navier-Stokes contexts: DGX_ERR_UNSUPPORTED.  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (other streams or host threads) race. */
dgx_status dgx_ssp_rk3_step_noalloc(dgx_ctx* ctx, double* u, double dt, int enforce_positivity, double* max_abs_r);

/* HIP graphs of the launch sequences above, recorded once and replayed with one hipGraphLaunch each: the
 * explicit time loop (solve_explicit.c:78-118 calling time_step_ssp_rk_33 every step) and repeated residual
 * evaluations then cost no per-kernel launch overhead.  A graph binds the pointers and dt it was recorded
 * with; it replays on the context's stream (dgx_set_stream) as of dgx_graph_launch. */
typedef struct dgx_graph dgx_graph;
enum { DGX_GRAPH_RHS_FP64 = 0, DGX_GRAPH_RHS_FP32 = 1, DGX_GRAPH_RHS_MIXED = 2 };
/* one residual evaluation: kind DGX_GRAPH_RHS_FP64 (dgx_compute_rhs: double u, r), _FP32 (dgx_compute_rhs_f32:
 * float u, r) or _MIXED (dgx_compute_rhs_mixed: double u, float r) */
dgx_status dgx_graph_create_rhs(dgx_ctx* ctx, int kind, const void* u, void* r, dgx_graph** g);
/* one dgx_ssp_rk3_step(ctx, u, dt, enforce_positivity, ...) */
dgx_status dgx_graph_create_rk3(dgx_ctx* ctx, double* u, double dt, int enforce_positivity, dgx_graph** g);
/* n replays, asynchronous (n >= 0) */
dgx_status dgx_graph_launch(dgx_graph* g, int n);
/* waits for the replays; for an RK3 graph *max_abs_r (HOST, may be NULL) receives max|M^-1 R| of the last stage
 * of the last replay, and DGX_ERR_POSITIVITY is returned if a limiter failed in any replay since the previous
 * dgx_graph_sync (dgx_ssp_rk3_step and dgx_enforce_positivity on the same context clear that flag too) */
dgx_status dgx_graph_sync(dgx_graph* g, double* max_abs_r);
dgx_status dgx_graph_destroy(dgx_graph* g);
/* Launch shape of the residual kernels.  The grid is either one workgroup per element block (the default,
 * grid 0) or persistent: `grid` workgroups, each XCD's share striding over that XCD's contiguous range of
 * element blocks with the reference-element operators staged in LDS once.  kind: DGX_GRAPH_RHS_FP64, _FP32,
 * _MIXED, or DGX_TUNE_RHS_MINV (the M^-1 R launch of the explicit step).  The persistent kernel is a separate
 * compilation of the same code; its results agree with the default kernel's to rounding (the compiler may
 * pair multiplies and adds into FMAs differently), not bit for bit. */
enum { DGX_TUNE_RHS_MINV = 3 };
dgx_status dgx_set_rhs_grid(dgx_ctx* ctx, int kind, int grid);
/* Autotuner: times the kernel of `kind` on (u, r) (device buffers of that kernel's types) for grid 0 and for
 * persistent grids of 1, 2, 3, 4, 6, 8, 12 and 16 workgroups per CU, keeps the fastest for this context and
 * mesh (a persistent grid only when more than 1% faster) and returns it in *grid and its time per launch (ms)
 * in *ms (HOST, either may be NULL).  Later launches and graph recordings of that kernel use it.  Synchronous;
 * not during a graph capture. */
dgx_status dgx_autotune_rhs(dgx_ctx* ctx, int kind, const void* u, void* r, int* grid, double* ms);
/* max |x|_inf over n doubles (HOST result). Synchronous. */
dgx_status dgx_max_abs(dgx_ctx* ctx, const double* x, long long n, double* out);
/* FP64 <-> FP32 state conversion on the device. */
dgx_status dgx_convert_f64_f32(dgx_ctx* ctx, const double* src, float* dst, long long n);
dgx_status dgx_convert_f32_f64(dgx_ctx* ctx, const float* src, double* dst, long long n);

/* Implicit path (contexts of one order: 2-D QUAD advection, Euler, and Navier-Stokes where stated; 3-D HEX
 * advection and Euler for J v, the Newton operator, the diagonal and face blocks, the block-sparse export, the
 * pseudo-time step and the solve with block-Jacobi up to nb = 143 or block SGS up to nb = 142 (both p <= 2), or
 * no preconditioner).
 * Diagonal blocks A_e = dR_e/du_e - mass_scale[e] * M_e (mass_scale: device array of E, or NULL for none), (Np*nv)^2 each, row-major, row eq*Np+node,
 * column var*Np+node: volume lhs_v_1 + face lhs_ll/lhs_rr contributions + CFL term
 * (constructor_lhs_v_1_T compute_volume_rlhs_T.c:179-220, constructor_lhs_f_1_T
 *  compute_face_rlhs_T.c:163-220, compute_CFL_ramping solve_dg.c:223-248 with mass_scale = 1/dt).
 * Navier-Stokes: the exact dR_e/du_e of the BR2 residual, the element's own lifted gradients differentiated with
 * the rest and the neighbours' states and gradient traces held fixed (the 'i' branch of compute_rlhs_dg,
 * compute_face_rlhs_dg_T.c:243-300). */
dgx_status dgx_assemble_diag_blocks(dgx_ctx* ctx, const double* u, const double* mass_scale, double* blocks);
/* Face (off-diagonal) blocks: blocks[(e * 2d + lf) * nb^2 + row * nb + col] = dR_e[row] / du_n[col], n the
 * neighbour across local face lf (col in n's local numbering var*Np+node); zero for boundary faces.  Only the
 * nv*N^(d-1) face-node rows and columns are non-zero (the neighbour enters through its trace, the face integral
 * tests with the element's face nodes); they are computed compactly, nv*N^(d-1) dual-number launches, and
 * scattered.
 * These are lhs_lr / lhs_rl of constructor_lhs_f_1_T (compute_face_rlhs_T.c:163-220, added at
 * compute_face_rlhs_dg.c:198-213); with the diagonal blocks they are the whole block-sparse Jacobian. */
/* Navier-Stokes (2-D): every row and column of a face block is non-zero (the neighbour's whole state enters the BR2
 * residual through its lifted gradient, compute_volume_rlhs_dg.c:220-261); nb dual-number launches. */
dgx_status dgx_assemble_face_blocks(dgx_ctx* ctx, const double* u, double* blocks);
/* The Newton matrix dR/du - mass_scale[e] M_e (mass_scale NULL: dR/du) in block-sparse row form, the sparsity of
 * the reference's AIJ matrix (constructor_nnz_dg_T, solve_dg_T.c:41-74; filled by add_to_petsc_Mat, solve.c:321-338):
 * block row e holds the element's own block and one block per face neighbour, columns ascending, nb = nv Np.
 * dgx_jacobian_bsr_pattern: host arrays row_ptr[E + 1] and col_idx[nnzb] (either may be NULL to query nnzb).
 * dgx_assemble_jacobian_bsr: device values[nnzb][nb][nb], row-major blocks in the pattern's order (2-D, and 3-D
 * advection and Euler; the diagonal and face blocks above, a neighbour reached through two faces getting their
 * sum); synchronous. */
dgx_status dgx_jacobian_bsr_pattern(dgx_ctx* ctx, int* row_ptr, int* col_idx, long long* nnzb);
dgx_status dgx_assemble_jacobian_bsr(dgx_ctx* ctx, const double* u, const double* mass_scale, double* values);
/* CFL-ramped pseudo-time step per element, returned as inv_dt[e] = max wave speed / (cfl * h_min), and for
 * Navier-Stokes the larger of that and mu / (cfl * h_min^2) (compute_dt_cfl_constrained /
 * compute_min_length_measure, solve_dg.c:284-379). */
dgx_status dgx_cfl_dt(dgx_ctx* ctx, const double* u, double cfl, double* inv_dt);
/* y_e -= scale[e] * M_e x_e (scale NULL = 1): the -M/dt term of the CFL-ramped system. */
dgx_status dgx_mass_sub(dgx_ctx* ctx, const double* x, const double* scale, double* y);
/* Exact Jacobian-vector product jv = (dR/du)(u) v by forward-mode (dual-number) evaluation of the
 * residual (Navier-Stokes: of the BR2 residual, gradients included): the matrix-free operator of the Newton
 * system (replaces the PETSc AIJ Mat assembled by add_to_petsc_Mat, solve.c:321-338). */
dgx_status dgx_jacobian_vector(dgx_ctx* ctx, const double* u, const double* v, double* jv);
/* out = (dR/du)(u) v - mass_scale[e] M_e v_e (mass_scale: device, E doubles): the operator of the pseudo-time
 * Newton system, J v and the -M/dt term in one dual-number launch (the GMRES operator of dgx_solve_implicit). */
dgx_status dgx_newton_operator(dgx_ctx* ctx, const double* u, const double* mass_scale, const double* v, double* out);
/* Batched LU (partial pivoting) of E blocks of size nb <= 143 in place (a block in LDS); piv: E*nb ints
 * (device). */
dgx_status dgx_block_lu_factor(dgx_ctx* ctx, double* blocks, int* piv, int E, int nb);
/* x <- A_e^{-1} x for every element block (x: E*nb). */
dgx_status dgx_block_lu_solve(dgx_ctx* ctx, const double* lu, const int* piv, double* x, int E, int nb);
/* In-place inverse of E blocks of size nb <= 142 (Gauss-Jordan, partial pivoting, a block in LDS), stored transposed:
 * blocks[e][j * nb + i] = (A_e^-1)[i][j].  The block-Jacobi preconditioner applies it as one mat-vec. */
dgx_status dgx_block_inverse(dgx_ctx* ctx, double* blocks, int E, int nb);
/* y_e = A_e^-1 x_e from the transposed inverses of dgx_block_inverse (x, y: E*nb, distinct). */
dgx_status dgx_block_inv_apply(dgx_ctx* ctx, const double* inv, const double* x, double* y, int E, int nb);

/* Steady solve by pseudo-time Newton-GMRES (solve_implicit, solve_implicit.c:95-110, 232-281, 393-470).
 * Names mirror the reference's test_case.data keys (cfl_initial, exit_tol_i, exit_ratio_i) and the PETSc
 * options file (GMRES restart 60, rtol 1e-4).  u (device) is updated in place.  Navier-Stokes contexts: the
 * FP64 or FP32 operator, any preconditioner but DGX_PC_LEARNED (block SGS over the full BR2 face blocks,
 * nb <= 142). */
/* Preconditioners of the GMRES solves.  DGX_PC_BLOCK_JACOBI: exact inverses of the element diagonal blocks of
 * dR/du - M/dt.  DGX_PC_BLOCK_SGS (nb <= 142: QUAD Euler p <= 4, HEX p <= 2): one forward and one backward multicolour block Gauss-Seidel sweep
 * over the diagonal and face blocks from z = 0, the coupled-block counterpart of the reference's default
 * PETSc ILU(1) with RCM ordering (external/petsc/options_files/petsc_options_gmres_default.txt:6-8), which
 * factors the assembled matrix sequentially.  DGX_PC_BLOCK_SGS32: the same sweeps with the inverse diagonal
 * blocks and the face blocks stored in FP32 (half the bytes the sweeps read; products and sums in FP64, the
 * vectors FP64).  DGX_PC_LEARNED: the MLP below. */
enum { DGX_PC_NONE = 0, DGX_PC_BLOCK_JACOBI = 1, DGX_PC_LEARNED = 2, DGX_PC_BLOCK_SGS = 3, DGX_PC_BLOCK_SGS32 = 4 };
/* z = P^-1 v for one preconditioner set up at state u with the pseudo-time scale inv_dt (E doubles, as
 * dgx_cfl_dt writes it): the operator that dgx_solve_implicit applies inside GMRES, exposed for callers that
 * run their own Krylov solver.  All pointers device, v and z distinct; synchronous. */
dgx_status dgx_precond_apply(dgx_ctx* ctx, int precond, const double* u, const double* inv_dt, const double* v,
                             double* z);
#define DGX_HISTORY_MAX 256
typedef struct dgx_implicit_params {
	double cfl_initial;   /* cfl_initial */
	double exit_tol;      /* exit_tol_i: stop when max|R| < exit_tol */
	double exit_ratio;    /* exit_ratio_i: stop when max|R|/max|R_0| < exit_ratio (0 = off) */
	int max_newton;
	int gmres_restart;    /* -ksp_gmres_restart */
	double gmres_rtol;    /* -ksp_rtol */
	int gmres_max_iter;   /* per Newton step; hitting it is DIVERGED_ITS: counted, Newton continues */
	int precond;          /* DGX_PC_* (DGX_PC_LEARNED needs dgx_attach_mlp) */
	double cfl_max;       /* upper bound of the ramped CFL number (0 = none, as the reference) */
	int learned_fallback_iters; /* DGX_PC_LEARNED: when GMRES has not converged after this many iterations the
	                               step is solved again with the exact block-Jacobi preconditioner (0 = never) */
	/* optional capture of the training set along the iterations (DGX_PC_BLOCK_JACOBI only): per Newton
	 * step the E diagonal blocks (E x nb x nb doubles) and features (E x nf floats) are appended while
	 * capacity (in elements) lasts; device pointers or NULL */
	double* capture_blocks;
	float* capture_features;   /* nf = round_up(nb + Np + 1, 16) floats per element */
	float* capture_sigma;      /* 1 float per element */
	long long capture_capacity;
	int jvp_fp32;         /* 1: the GMRES operator dR/du v - M v / dt in FP32 dual numbers (FP64 vectors; Euler:
	                         FP32 geometry and operators, Navier-Stokes: FP64 ones converted as they are read): the
	                         inner defect-correction product; the FP64 residual of
	                         each Newton step still decides convergence.  0: FP64 (no counterpart: the reference
	                         assembles the FP64 Jacobian) */
} dgx_implicit_params;
typedef struct dgx_implicit_stats {
	int newton_steps;
	int gmres_iters;
	int gmres_not_converged;
	int converged;
	double max_rhs0;
	double max_rhs;
	double seconds;
	long long captured;   /* elements appended to the capture buffers */
	int fallback_steps;   /* Newton steps re-solved with block-Jacobi (learned_fallback_iters) */
	int gmres_per_step[DGX_HISTORY_MAX];   /* GMRES iterations of each Newton step (all attempts) */
	double history[DGX_HISTORY_MAX];
} dgx_implicit_stats;
dgx_status dgx_solve_implicit(dgx_ctx* ctx, double* u, const dgx_implicit_params* params, dgx_implicit_stats* stats);

/* Learned block preconditioner (no counterpart in the reference; the reference's preconditioner is PETSc
 * ILU(1)).  A 3-layer MLP with BF16 weights and FP32 accumulation maps element features
 * [state / u_ref (nb) | det J / mean (Np) | log10(h / (dt * wave speed))] to an approximate inverse of the
 * element diagonal block, scaled by sigma_e = 1 / (mean(w det J)/dt_e + 0.25 h_e lambda_e):
 *   DGX_MLP_LOWRANK: P_e = sigma_e (diag(d_e) + U_e V_e^T), rank 8 (outputs nb + 2 nb 8);
 *   DGX_MLP_KRON:    P_e = sigma_e sum_{r<R} S_r (x) Q_r, S_r nv x nv, Q_r Np x Np, R = 4 or 8
 *                    (outputs R (nv^2 + Np^2)).
 * Both start from P_e = -sigma_e I.  2-D Euler with p = 2 (nb 36) or p = 3 (nb 64).
 * Trained with its own forward/backward kernels on the solver's diagonal blocks. */
enum { DGX_MLP_LOWRANK = 0, DGX_MLP_KRON = 1 };
dgx_status dgx_mlp_create(dgx_mlp** m, int kind, int nb, int Np, int hidden, int rank, unsigned long long seed);
dgx_status dgx_mlp_destroy(dgx_mlp* m);
/* nf: feature count (padded to 16), nout: outputs per element, nparam: FP32 master parameter count */
dgx_status dgx_mlp_sizes(const dgx_mlp* m, int* nf, int* nout, long long* nparam);
dgx_status dgx_mlp_get_params(dgx_mlp* m, float* host);        /* FP32 master weights W1,b1,W2,b2,W3,b3 */
dgx_status dgx_mlp_set_params(dgx_mlp* m, const float* host);
/* features (E x nf floats, device) and the per-element scale sigma (E floats, device) for state u and inv_dt */
dgx_status dgx_mlp_features(dgx_ctx* ctx, const dgx_mlp* m, const double* u, const double* inv_dt, float* feat,
                            float* sigma);
/* mean loss and (optionally) the parameter gradient (HOST, nparam floats) on B samples (device inputs;
 * sigma may be NULL = 1): the quantities the training step uses, exposed for gradient checks */
dgx_status dgx_mlp_loss_grad(dgx_mlp* m, const float* feat, const float* sigma, const double* blocks, int B,
                             double* loss, float* grad_host);
/* Adam on n samples in mini-batches of `batch` for `epochs` passes; *loss receives the last full-set loss */
dgx_status dgx_mlp_train(dgx_mlp* m, const float* feat, const float* sigma, const double* blocks, long long n,
                         int epochs, int batch, float lr, double* loss);
/* outputs (B x nout floats, device) for given features */
dgx_status dgx_mlp_predict(dgx_mlp* m, const float* feat, int B, float* out);
/* use m as the DGX_PC_LEARNED preconditioner of ctx (m must outlive the solves) */
dgx_status dgx_attach_mlp(dgx_ctx* ctx, dgx_mlp* m);

#ifdef __cplusplus
}
#endif
#endif
