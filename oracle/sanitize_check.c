/*
 * sanitize_check.c -- TEST INFRASTRUCTURE ONLY: runs the oracle (dg_oracle.c) under the address and
 * undefined-behaviour sanitizers (`make -C oracle sanitize_check`, then oracle/_san/sanitize_check).
 *
 * Small periodic QUAD (p = 1..4) and HEX (p = 1..3) meshes, slightly curved, with a smooth Euler state: the
 * residual (sequential, threaded, volume and face parts, complex twin), M^-1, the explicit residual, the
 * positivity limiter, an SSP-RK3 step, the complex-step blocks, the BR2 gradient and Navier-Stokes residual
 * (2-D), and the point functions (fluxes, boundary states) on a spread of states.  Exit status 0 when every
 * result is finite and the residual's parts add up; the sanitizers abort on any memory or undefined-behaviour
 * fault.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dg_oracle.h"

static int failures = 0;

static void expect(int ok, const char* what, int d, int p)
{
	if (!ok) {
		fprintf(stderr, "sanitize_check: %s failed (d=%d p=%d)\n", what, d, p);
		++failures;
	}
}

static int finite_all(const double* x, size_t n)
{
	for (size_t i = 0; i < n; ++i)
		if (!isfinite(x[i])) return 0;
	return 1;
}

/* n^d periodic tensor-product mesh of [0,1]^d, curved a little, faces {eL, fL, eR, fR, ord, bc} */
static void tp_mesh(int d, int p, int n, double** xyz, int** faces, int* E, int* F)
{
	const int N = p + 1;
	int Np = 1, nE = 1;
	for (int k = 0; k < d; ++k) {
		Np *= N;
		nE *= n;
	}
	double x1[8], w1[8];
	or_gauss_lobatto(N, x1, w1);
	*E = nE;
	*F = d * nE;
	*xyz = (double*)malloc(sizeof(double) * (size_t)nE * d * Np);
	*faces = (int*)malloc(sizeof(int) * 6 * (size_t)(*F));
	for (int e = 0; e < nE; ++e) {
		int ij[3] = { e % n, (e / n) % n, d == 3 ? e / (n * n) : 0 };
		for (int a = 0; a < Np; ++a) {
			const int ab[3] = { a % N, (a / N) % N, d == 3 ? a / (N * N) : 0 };
			double x[3];
			for (int k = 0; k < d; ++k) x[k] = (ij[k] + 0.5 * (x1[ab[k]] + 1.0)) / n;
			for (int k = 0; k < d; ++k) {
				const double o = x[(k + 1) % d];
				(*xyz)[((size_t)e * d + k) * Np + a] = x[k] + 0.02 * sin(2.0 * M_PI * o);
			}
		}
		for (int k = 0; k < d; ++k) {
			int jk[3] = { ij[0], ij[1], ij[2] };
			jk[k] = (jk[k] + 1) % n;
			const int en = jk[0] + n * (jk[1] + n * jk[2]);
			int* f = *faces + 6 * ((size_t)e * d + k);
			f[0] = e;
			f[1] = 2 * k + 1;
			f[2] = en;
			f[3] = 2 * k;
			f[4] = 0;
			f[5] = OR_BC_INTERIOR;
		}
	}
}

static void smooth_state(int d, int Np, int E, const double* xyz, double* u)
{
	const int nv = d + 2;
	for (int e = 0; e < E; ++e)
		for (int a = 0; a < Np; ++a) {
			double x[3] = { 0, 0, 0 };
			for (int k = 0; k < d; ++k) x[k] = xyz[((size_t)e * d + k) * Np + a];
			const double rho = 1.0 + 0.1 * sin(2 * M_PI * x[0]) * cos(2 * M_PI * x[1]);
			const double vel[3] = { 0.3, -0.2, 0.1 };
			const double pr = 1.0 + 0.05 * cos(2 * M_PI * (x[0] + x[1] + x[2]));
			double ke = 0.0;
			double* s = u + (size_t)e * nv * Np + a;
			s[0] = rho;
			for (int k = 0; k < d; ++k) {
				s[(1 + k) * Np] = rho * vel[k];
				ke += vel[k] * vel[k];
			}
			s[(d + 1) * Np] = pr / OR_GM1 + 0.5 * rho * ke;
		}
}

static void check_mesh(int d, int p, int flux)
{
	const int n = 3;
	double* xyz;
	int* faces;
	int E, F;
	tp_mesh(d, p, n, &xyz, &faces, &E, &F);
	const double b_adv[3] = { 1.0, 0.5, 0.25 }, bc_prim[5] = { 1.0, 0.3, 0.0, 0.0, 1.0 };
	or_ctx* c = or_create(d, p, OR_PDE_EULER, flux, E, F, xyz, faces, b_adv, bc_prim);
	expect(c != NULL, "or_create", d, p);
	if (!c) {
		free(xyz);
		free(faces);
		return;
	}
	const int Np = or_np(c), nv = or_nv(c);
	const size_t ndof = (size_t)E * Np * nv;
	double* u = (double*)malloc(sizeof(double) * ndof);
	double* r = (double*)malloc(sizeof(double) * ndof);
	double* rv = (double*)malloc(sizeof(double) * ndof);
	double* rf = (double*)malloc(sizeof(double) * ndof);
	double* rt = (double*)malloc(sizeof(double) * ndof);
	double* zi = (double*)calloc(ndof, sizeof(double));
	double* ri = (double*)malloc(sizeof(double) * ndof);
	smooth_state(d, Np, E, xyz, u);
	or_compute_rhs_d(c, u, r);
	expect(finite_all(r, ndof), "residual finite", d, p);
	memset(rv, 0, sizeof(double) * ndof);
	memset(rf, 0, sizeof(double) * ndof);
	or_volume_rhs_d(c, u, rv);
	or_face_rhs_d(c, u, rf);
	double worst = 0.0, scale = 0.0;
	for (size_t i = 0; i < ndof; ++i) {
		worst = fmax(worst, fabs(rv[i] + rf[i] - r[i]));
		scale = fmax(scale, fabs(rv[i]));
	}
	expect(worst <= 1e-12 * (1.0 + scale), "volume + face = residual", d, p);
	or_compute_rhs_d_threads(c, u, rt, 3);
	worst = 0.0;
	for (size_t i = 0; i < ndof; ++i) worst = fmax(worst, fabs(rt[i] - r[i]));
	expect(worst <= 1e-12 * (1.0 + scale), "threaded residual", d, p);
	or_compute_rhs_c(c, u, zi, rt, ri);
	expect(finite_all(rt, ndof) && finite_all(ri, ndof), "complex residual finite", d, p);
	memcpy(rt, r, sizeof(double) * ndof);
	expect(or_apply_minv(c, rt) == 0 && finite_all(rt, ndof), "M^-1", d, p);
	expect(isfinite(or_compute_rhs_explicit(c, u, rt)), "explicit residual", d, p);
	expect(or_enforce_positivity(c, u) == 0, "positivity limiter", d, p);
	int err = 0;
	const double mx = or_ssp_rk3_step(c, u, 1e-4, 1, &err);
	expect(err == 0 && isfinite(mx) && finite_all(u, ndof), "SSP-RK3 step", d, p);
	if (d == 2) {
		const size_t nb = (size_t)Np * nv;
		double* blocks = (double*)malloc(sizeof(double) * (size_t)E * nb * nb);
		or_diag_blocks_cs(c, u, 1e-30, blocks);
		expect(finite_all(blocks, (size_t)E * nb * nb), "complex-step blocks", d, p);
		or_jacobian_block_cs(c, u, 0, 1, 1e-30, blocks);
		expect(finite_all(blocks, nb * nb), "complex-step face block", d, p);
		free(blocks);
		or_set_viscous(c, 0.01, 0.72);
		const size_t ng = (size_t)E * d * nv * Np;
		double* g = (double*)malloc(sizeof(double) * ng);
		double* gv = (double*)malloc(sizeof(double) * ng);
		double* gf = (double*)malloc(sizeof(double) * (size_t)F * 2 * d * nv * Np);
		expect(or_ns_gradient(c, u, g, gv, gf) == 0 && finite_all(g, ng), "BR2 gradient", d, p);
		expect(or_compute_rhs_ns(c, u, rt) == 0 && finite_all(rt, ndof), "Navier-Stokes residual", d, p);
		free(g);
		free(gv);
		free(gf);
	}
	or_destroy(c);
	free(u);
	free(r);
	free(rv);
	free(rf);
	free(rt);
	free(zi);
	free(ri);
	free(xyz);
	free(faces);
}

static void check_points(void)
{
	for (int d = 2; d <= 3; ++d) {
		const int nv = d + 2;
		for (int k = 0; k < 64; ++k) {
			double sL[5], sR[5], n[3] = { 0, 0, 0 }, f[15], nnf[5], sb[5];
			const double th = 0.37 * k, ph = 0.11 * k;
			n[0] = cos(th);
			n[1] = sin(th) * (d == 3 ? cos(ph) : 1.0);
			if (d == 3) n[2] = sin(th) * sin(ph);
			for (int side = 0; side < 2; ++side) {
				double* s = side ? sR : sL;
				const double rho = 0.5 + 0.01 * ((k * 7 + side * 3) % 50);
				const double M = 0.05 * ((k * 5 + side) % 40);   /* Mach 0 .. 2: every Riemann BC branch */
				const double pr = 0.6 + 0.01 * ((k * 3 + side * 11) % 60);
				const double cs = sqrt(OR_GAMMA * pr / rho);
				double ke = 0.0;
				s[0] = rho;
				for (int j = 0; j < d; ++j) {
					const double v = M * cs * (j == 0 ? n[0] : j == 1 ? -n[1] : 0.3);
					s[1 + j] = rho * v;
					ke += v * v;
				}
				s[d + 1] = pr / OR_GM1 + 0.5 * rho * ke;
			}
			or_euler_flux_d(d, sL, f);
			for (int fl = OR_FLUX_LF; fl <= OR_FLUX_ROE; ++fl) {
				or_numflux_d(d, OR_PDE_EULER, fl, sL, sR, n, NULL, nnf);
				expect(finite_all(nnf, (size_t)nv), "numerical flux", d, fl);
			}
			const double prim[5] = { 1.0, 0.4, 0.1, d == 3 ? 0.0 : 1.0, 1.0 };
			for (int bc = OR_BC_RIEMANN; bc <= OR_BC_SLIPWALL; ++bc) {
				or_bc_state_d(d, bc, sL, n, prim, sb);
				expect(finite_all(sb, (size_t)nv), "boundary state", d, bc);
			}
		}
	}
}

int main(void)
{
	check_points();
	for (int p = 1; p <= 4; ++p) check_mesh(2, p, p % 2 ? OR_FLUX_ROE : OR_FLUX_LF);
	for (int p = 1; p <= 3; ++p) check_mesh(3, p, p % 2 ? OR_FLUX_LF : OR_FLUX_ROE);
	if (failures) return 1;
	printf("sanitize_check: ok\n");
	return 0;
}
