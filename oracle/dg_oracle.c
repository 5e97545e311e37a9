/*
 * dg_oracle.c -- CPU FP64 restatement of DPGSolver's DG residual path.
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path, never part of it.
 * See dg_oracle.h for the list of reference routines restated here.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "dg_oracle.h"

struct or_ctx {
	int d, p, N, Np, Nc, Nfc, nv, pde, flux, E, F, n_ord;
	int* faces;
	double b_adv[3];
	double bc_prim[5];
	double xgll[16], wgll[16], xgl[16], wgl[16];
	double* cv0_vs_vc;   /* Nc x Np */
	double* cv1_vs_vc;   /* d x (Nc x Np) */
	double* tw1_vt_vc;   /* d x (Np x Nc) */
	double* cv0_vs_fc;   /* 2d x (Nfc x Np) */
	double* tw0_vt_fc;   /* 2d x (Np x Nfc) */
	double* d_vm;        /* d x (Np x Np): GLL differentiation on the metric nodes */
	double* w_vc;        /* Nc */
	double* w_fc;        /* Nfc */
	int* perm;           /* n_ord x Nfc */
	double* xyz;         /* E x d x Np */
	double* metrics_vm;  /* E x d*d x Np */
	double* metrics_vc;  /* E x d*d x Nc */
	double* detJ_vc;     /* E x Nc */
	double* nJ_fc;       /* F x Nfc x d */
	double* minv;        /* E x Np x Np (lazy) */
	double* ccBS;        /* Np x Np: Bezier -> nodal */
	double* ccSB;        /* Np x Np: nodal -> Bezier */
	int* ef_off;         /* E+1 */
	int* ef_list;        /* (face << 1) | side */
	int viscous;         /* Navier-Stokes: Euler + BR2 viscous terms (or_set_viscous) */
	double mu, Pr;       /* constant viscosity (compute_mu_constant_T) and Prandtl number */
	double wall_nfE;     /* no-slip flux walls: the normal viscous energy flux nf_E (0 adiabatic) */
	double wall_omega;   /* rotating no-slip walls: angular velocity, and the wall density and total energy */
	double wall_rho, wall_E;
	int sutherland;      /* viscosity by Sutherland's formula (or_set_sutherland) */
	double r_s;
};

/* ------------------------------------------------------------------------------------------------ */
/* 1-D nodes (nodes.c:69-104 uses the Jacobi-polynomial package; restated as Newton on Legendre)    */

static void legendre(int n, double x, double* P, double* dP)
{
	/* P_n and P_n' by the three-term recurrence */
	double p0 = 1.0, p1 = x;
	if (n == 0) { *P = 1.0; *dP = 0.0; return; }
	for (int k = 2; k <= n; ++k) {
		const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
		p0 = p1;
		p1 = p2;
	}
	*P = p1;
	*dP = (fabs(x) < 1.0) ? n * (x * p1 - p0) / (x * x - 1.0) : 0.5 * n * (n + 1.0) * pow(x, n + 1);
}

int or_gauss_legendre(int n, double* x, double* w)
{
	if (n < 1 || n > 16) return -1;
	for (int i = 0; i < n; ++i) {
		double z = -cos(M_PI * (i + 0.75) / (n + 0.5));
		for (int it = 0; it < 100; ++it) {
			double P, dP;
			legendre(n, z, &P, &dP);
			const double dz = P / dP;
			z -= dz;
			if (fabs(dz) < 1e-16) break;
		}
		double P, dP;
		legendre(n, z, &P, &dP);
		x[i] = z;
		w[i] = 2.0 / ((1.0 - z * z) * dP * dP);
	}
	/* exact symmetry */
	for (int i = 0; i < n / 2; ++i) {
		const double a = 0.5 * (x[n - 1 - i] - x[i]);
		x[i] = -a; x[n - 1 - i] = a;
		const double b = 0.5 * (w[i] + w[n - 1 - i]);
		w[i] = b; w[n - 1 - i] = b;
	}
	if (n % 2) x[n / 2] = 0.0;
	return 0;
}

int or_gauss_lobatto(int n, double* x, double* w)
{
	if (n < 2 || n > 16) return -1;
	const int N = n - 1; /* interior nodes are roots of P_N' */
	x[0] = -1.0; x[n - 1] = 1.0;
	for (int i = 1; i < N; ++i) {
		double z = -cos(M_PI * i / N);
		for (int it = 0; it < 100; ++it) {
			/* Newton on (1-z^2) P_N'(z) via q = P_{N+1} - P_{N-1} */
			double P, dP, Pm, dPm, Pp, dPp;
			legendre(N, z, &P, &dP);
			legendre(N - 1, z, &Pm, &dPm);
			legendre(N + 1, z, &Pp, &dPp);
			const double q = Pp - Pm, dq = dPp - dPm;
			const double dz = q / dq;
			z -= dz;
			if (fabs(dz) < 1e-16) break;
		}
		x[i] = z;
	}
	for (int i = 0; i < n; ++i) {
		double P, dP;
		legendre(N, x[i], &P, &dP);
		w[i] = 2.0 / (N * (N + 1.0) * P * P);
	}
	for (int i = 0; i < n / 2; ++i) {
		const double a = 0.5 * (x[n - 1 - i] - x[i]);
		x[i] = -a; x[n - 1 - i] = a;
		const double b = 0.5 * (w[i] + w[n - 1 - i]);
		w[i] = b; w[n - 1 - i] = b;
	}
	if (n % 2) x[n / 2] = 0.0;
	return 0;
}

void or_lagrange(int nin, const double* xin, int nout, const double* xout, double* Iop, double* Dop)
{
	for (int r = 0; r < nout; ++r) {
		const double x = xout[r];
		for (int c = 0; c < nin; ++c) {
			double l = 1.0, den = 1.0;
			for (int k = 0; k < nin; ++k)
				if (k != c) { l *= (x - xin[k]); den *= (xin[c] - xin[k]); }
			if (Iop) Iop[r * nin + c] = l / den;
			if (Dop) {
				double s = 0.0;
				for (int j = 0; j < nin; ++j) {
					if (j == c) continue;
					double t = 1.0;
					for (int k = 0; k < nin; ++k)
						if (k != c && k != j) t *= (x - xin[k]);
					s += t;
				}
				Dop[r * nin + c] = s / den;
			}
		}
	}
}

void or_kron(int d, const double* const* ops, const int* rows, const int* cols, double* out)
{
	int R = 1, C = 1;
	for (int k = 0; k < d; ++k) { R *= rows[k]; C *= cols[k]; }
	for (int r = 0; r < R; ++r)
		for (int c = 0; c < C; ++c) {
			int rr = r, cc = c;
			double v = 1.0;
			for (int k = 0; k < d; ++k) {
				const int ir = rr % rows[k], ic = cc % cols[k];
				rr /= rows[k]; cc /= cols[k];
				v *= ops[k][ir * cols[k] + ic];
			}
			out[(size_t)r * C + c] = v;
		}
}

void or_bernstein(int p, int n, const double* x, double* B)
{
	for (int r = 0; r < n; ++r) {
		const double a = 0.5 * (1.0 + x[r]), b = 0.5 * (1.0 - x[r]);
		for (int j = 0; j <= p; ++j) {
			double binom = 1.0;
			for (int k = 1; k <= j; ++k) binom = binom * (p - j + k) / k;
			B[r * (p + 1) + j] = binom * pow(a, j) * pow(b, p - j);
		}
	}
}

/* ------------------------------------------------------------------------------------------------ */
/* small dense linear algebra                                                                        */

static int invert(int n, const double* A, double* Ainv)
{
	double* M = (double*)malloc(sizeof(double) * n * 2 * n);
	for (int i = 0; i < n; ++i)
		for (int j = 0; j < 2 * n; ++j)
			M[i * 2 * n + j] = (j < n ? A[i * n + j] : (j - n == i ? 1.0 : 0.0));
	for (int col = 0; col < n; ++col) {
		int piv = col;
		for (int i = col + 1; i < n; ++i)
			if (fabs(M[i * 2 * n + col]) > fabs(M[piv * 2 * n + col])) piv = i;
		if (M[piv * 2 * n + col] == 0.0) { free(M); return -1; }
		if (piv != col)
			for (int j = 0; j < 2 * n; ++j) {
				const double t = M[col * 2 * n + j];
				M[col * 2 * n + j] = M[piv * 2 * n + j];
				M[piv * 2 * n + j] = t;
			}
		const double inv = 1.0 / M[col * 2 * n + col];
		for (int j = 0; j < 2 * n; ++j) M[col * 2 * n + j] *= inv;
		for (int i = 0; i < n; ++i) {
			if (i == col) continue;
			const double f = M[i * 2 * n + col];
			if (f == 0.0) continue;
			for (int j = 0; j < 2 * n; ++j) M[i * 2 * n + j] -= f * M[col * 2 * n + j];
		}
	}
	for (int i = 0; i < n; ++i)
		for (int j = 0; j < n; ++j) Ainv[i * n + j] = M[i * 2 * n + n + j];
	free(M);
	return 0;
}

/* ------------------------------------------------------------------------------------------------ */
/* templated half                                                                                     */

/* real-only helpers of the templated Navier-Stokes part, defined below */
static int ensure_minv(or_ctx* c);
static void face_xyz(const or_ctx* c, int e, int lf, double* xy);
static int ns_bc_supported(int bc);

#define Type double
#define FN(name) name##_d
#define SQRT_T sqrt
#define POW_T pow
#define REAL_T(x) (x)
#include "dg_oracle_T.inc"
#undef Type
#undef FN
#undef SQRT_T
#undef POW_T
#undef REAL_T

#define Type double complex
#define FN(name) name##_c
#define SQRT_T csqrt
#define POW_T cpow
#define REAL_T(x) creal(x)
#include "dg_oracle_T.inc"
#undef Type
#undef FN
#undef SQRT_T
#undef POW_T
#undef REAL_T

/* ------------------------------------------------------------------------------------------------ */
/* face-local node ordering and orientation permutations                                              */

/* volume node index of face-local node (a, b) on local face lf; GLL endpoints carry the face */
static int face_vol_node(int d, int N, int lf, int a, int b)
{
	const int fd = lf / 2, side = lf % 2;
	int idx[3] = { 0, 0, 0 }, k = 0;
	const int ab[2] = { a, b };
	for (int dim = 0; dim < d; ++dim) {
		if (dim == fd) idx[dim] = side ? N - 1 : 0;
		else idx[dim] = ab[k++];
	}
	return idx[0] + N * (idx[1] + N * idx[2]);
}

/* Face node correspondence for the reference's orientation index ind_ord (constructor_face_corr,
 * src/element/nodes_correspondence.c; golden table input/testing/unit/nodes/nodes_correspondence.data):
 * perm[i] is the right-face node that coincides with left-face node i, so a right trace is taken into the
 * left ordering as s_l[i] = s_r[perm[i]] (permute_Multiarray_T_fc, compute_face_rlhs_T.c:312-318).
 * LINE faces (2-D): 0 aligned, 1 reversed.  QUAD faces (3-D), face node (a, b) with a fastest: cases 4-7 start
 * from the transposed numbering, cases 1, 3, 6, 7 reverse within each row of N (the first face coordinate)
 * and cases 2, 3, 5, 7 reverse the order of the rows (the second). */
void or_face_perm(const or_ctx* c, int ord, int* perm)
{
	const int N = c->N;
	if (c->d == 1) { perm[0] = 0; return; }
	if (c->d == 2) {
		for (int a = 0; a < N; ++a) perm[a] = (ord == 1) ? N - 1 - a : a;
		return;
	}
	const int transposed = (ord >= 4);
	const int rev_rows = (ord == 1 || ord == 3 || ord == 6 || ord == 7);
	const int rev_blocks = (ord == 2 || ord == 3 || ord == 5 || ord == 7);
	for (int b = 0; b < N; ++b)
		for (int a = 0; a < N; ++a) {
			const int a2 = rev_rows ? N - 1 - a : a, b2 = rev_blocks ? N - 1 - b : b;
			perm[a + N * b] = transposed ? (b2 + N * a2) : (a2 + N * b2);
		}
}

/* ------------------------------------------------------------------------------------------------ */
/* geometry: geometry_T.c:734-801 (2-D), deprecated/src/setup_geom_factors.c:157-243 (3-D curl form) */

static void compute_geometry(or_ctx* c)
{
	const int d = c->d, Np = c->Np, Nc = c->Nc, Nfc = c->Nfc;
	double* jac = (double*)malloc(sizeof(double) * d * d * Np);
	double* jac_vc = (double*)malloc(sizeof(double) * d * d * Nc);
	double* V = (double*)malloc(sizeof(double) * 3 * Np);
	double* curl1 = (double*)malloc(sizeof(double) * 3 * Np);
	double* curl2 = (double*)malloc(sizeof(double) * 3 * Np);
	for (int e = 0; e < c->E; ++e) {
		const double* X = c->xyz + (size_t)e * d * Np;
		double* mvm = c->metrics_vm + (size_t)e * d * d * Np;
		/* jac[(i*d + k)*Np + n] = dX_i/dxi_k at the GLL metric nodes */
		for (int i = 0; i < d; ++i)
			for (int k = 0; k < d; ++k)
				for (int n = 0; n < Np; ++n) {
					double acc = 0.0;
					for (int m = 0; m < Np; ++m) acc += c->d_vm[(size_t)k * Np * Np + n * Np + m] * X[i * Np + m];
					jac[(i * d + k) * Np + n] = acc;
				}
		if (d == 1) {
			for (int n = 0; n < Np; ++n) mvm[n] = 1.0;
		} else if (d == 2) {
			for (int n = 0; n < Np; ++n) {
				const double x_r = jac[0 * Np + n], x_s = jac[1 * Np + n];
				const double y_r = jac[2 * Np + n], y_s = jac[3 * Np + n];
				mvm[(0 + 2 * 0) * Np + n] = y_s;
				mvm[(0 + 2 * 1) * Np + n] = -y_r;
				mvm[(1 + 2 * 0) * Np + n] = -x_s;
				mvm[(1 + 2 * 1) * Np + n] = x_r;
			}
		} else {
			/* row (physical dim) -> ordered pair (a, b): C = 0.5*(curl(X_a grad X_b) - curl(X_b grad X_a)) */
			static const int pairs[3][2] = { { 1, 2 }, { 2, 0 }, { 0, 1 } };
			for (int row = 0; row < 3; ++row) {
				for (int t = 0; t < 2; ++t) {
					const int a = t == 0 ? pairs[row][0] : pairs[row][1];
					const int b = t == 0 ? pairs[row][1] : pairs[row][0];
					/* V_k = X_a * dX_b/dxi_k */
					for (int k = 0; k < 3; ++k)
						for (int n = 0; n < Np; ++n) V[k * Np + n] = X[a * Np + n] * jac[(b * 3 + k) * Np + n];
					double* cu = t == 0 ? curl1 : curl2;
					static const int cyc[3][2] = { { 1, 2 }, { 2, 0 }, { 0, 1 } };
					for (int col = 0; col < 3; ++col) {
						const int k1 = cyc[col][0], k2 = cyc[col][1];
						for (int n = 0; n < Np; ++n) {
							double s1 = 0.0, s2 = 0.0;
							for (int m = 0; m < Np; ++m) {
								s1 += c->d_vm[(size_t)k1 * Np * Np + n * Np + m] * V[k2 * Np + m];
								s2 += c->d_vm[(size_t)k2 * Np * Np + n * Np + m] * V[k1 * Np + m];
							}
							cu[col * Np + n] = s1 - s2;
						}
					}
				}
				for (int col = 0; col < 3; ++col)
					for (int n = 0; n < Np; ++n)
						mvm[(row + 3 * col) * Np + n] = 0.5 * (curl1[col * Np + n] - curl2[col * Np + n]);
			}
		}
		/* metrics_vc = vv0_vm_vc * metrics_vm */
		double* mvc = c->metrics_vc + (size_t)e * d * d * Nc;
		for (int q = 0; q < d * d; ++q)
			for (int n = 0; n < Nc; ++n) {
				double acc = 0.0;
				for (int m = 0; m < Np; ++m) acc += c->cv0_vs_vc[n * Np + m] * mvm[q * Np + m];
				mvc[q * Nc + n] = acc;
			}
		/* detJ at vc from the Jacobian at vc */
		for (int i = 0; i < d; ++i)
			for (int k = 0; k < d; ++k)
				for (int n = 0; n < Nc; ++n) {
					double acc = 0.0;
					for (int m = 0; m < Np; ++m)
						acc += c->cv1_vs_vc[(size_t)k * Nc * Np + n * Np + m] * X[i * Np + m];
					jac_vc[(i * d + k) * Nc + n] = acc;
				}
		double* dj = c->detJ_vc + (size_t)e * Nc;
		for (int n = 0; n < Nc; ++n) {
#define JV(i, k) jac_vc[((i) * d + (k)) * Nc + n]
			if (d == 1) dj[n] = JV(0, 0);
			else if (d == 2) dj[n] = JV(0, 0) * JV(1, 1) - JV(0, 1) * JV(1, 0);
			else
				dj[n] = JV(0, 0) * (JV(1, 1) * JV(2, 2) - JV(1, 2) * JV(2, 1))
				      - JV(0, 1) * (JV(1, 0) * JV(2, 2) - JV(1, 2) * JV(2, 0))
				      + JV(0, 2) * (JV(1, 0) * JV(2, 1) - JV(1, 1) * JV(2, 0));
#undef JV
		}
	}
	/* faces: metrics_fc from the LEFT volume, v_i = sum_j nref_j m[i + d j] (geometry_T.c:412-445) */
	for (int f = 0; f < c->F; ++f) {
		const int eL = c->faces[6 * f], fL = c->faces[6 * f + 1];
		const double* mvm = c->metrics_vm + (size_t)eL * d * d * Np;
		const double* op = c->cv0_vs_fc + (size_t)fL * Nfc * Np;
		const int fd = fL / 2;
		const double sgn = (fL % 2) ? 1.0 : -1.0;
		double* nJ = c->nJ_fc + (size_t)f * Nfc * d;
		for (int n = 0; n < Nfc; ++n)
			for (int i = 0; i < d; ++i) {
				/* only the fd-th reference normal component is nonzero for TP faces */
				double acc = 0.0;
				for (int m = 0; m < Np; ++m) acc += op[n * Np + m] * mvm[(i + d * fd) * Np + m];
				nJ[n * d + i] = sgn * acc;
			}
	}
	free(jac); free(jac_vc); free(V); free(curl1); free(curl2);
}

/* ------------------------------------------------------------------------------------------------ */

or_ctx* or_create(int d, int p, int pde, int flux, int E, int F, const double* xyz, const int* faces,
                  const double* b_adv, const double* bc_prim)
{
	if (d < 1 || d > 3 || p < 1 || p > 15) return NULL;
	or_ctx* c = (or_ctx*)calloc(1, sizeof(or_ctx));
	c->d = d; c->p = p; c->N = p + 1; c->pde = pde; c->flux = flux; c->E = E; c->F = F;
	const int N = c->N;
	c->Np = (d == 1) ? N : (d == 2 ? N * N : N * N * N);
	c->Nc = c->Np;
	c->Nfc = (d == 1) ? 1 : (d == 2 ? N : N * N);
	c->nv = (pde == OR_PDE_EULER) ? d + 2 : 1;
	for (int k = 0; k < 3; ++k) c->b_adv[k] = b_adv ? b_adv[k] : 0.0;
	for (int k = 0; k < 5; ++k) c->bc_prim[k] = bc_prim ? bc_prim[k] : 0.0;
	or_gauss_lobatto(N, c->xgll, c->wgll);
	or_gauss_legendre(N, c->xgl, c->wgl);

	const int Np = c->Np, Nc = c->Nc, Nfc = c->Nfc;
	double I1[256], D1[256], Dg[256], one0[16], oneN[16], Id[256];
	or_lagrange(N, c->xgll, N, c->xgl, I1, D1);
	or_lagrange(N, c->xgll, N, c->xgll, NULL, Dg);
	for (int i = 0; i < N * N; ++i) Id[i] = 0.0;
	for (int i = 0; i < N; ++i) { Id[i * N + i] = 1.0; one0[i] = (i == 0); oneN[i] = (i == N - 1); }

	const int rowsN[3] = { N, N, N }, rows1[3] = { 1, 1, 1 };
	(void)rows1;
	const double* ops[3];

	c->cv0_vs_vc = (double*)malloc(sizeof(double) * Nc * Np);
	for (int k = 0; k < d; ++k) ops[k] = I1;
	or_kron(d, ops, rowsN, rowsN, c->cv0_vs_vc);

	c->cv1_vs_vc = (double*)malloc(sizeof(double) * d * Nc * Np);
	c->tw1_vt_vc = (double*)malloc(sizeof(double) * d * Np * Nc);
	c->d_vm = (double*)malloc(sizeof(double) * d * Np * Np);
	c->w_vc = (double*)malloc(sizeof(double) * Nc);
	for (int n = 0; n < Nc; ++n) {
		int r = n;
		double w = 1.0;
		for (int k = 0; k < d; ++k) { w *= c->wgl[r % N]; r /= N; }
		c->w_vc[n] = w;
	}
	for (int dim = 0; dim < d; ++dim) {
		for (int k = 0; k < d; ++k) ops[k] = (k == dim) ? D1 : I1;
		double* cv1 = c->cv1_vs_vc + (size_t)dim * Nc * Np;
		or_kron(d, ops, rowsN, rowsN, cv1);
		double* tw1 = c->tw1_vt_vc + (size_t)dim * Np * Nc;
		for (int m = 0; m < Np; ++m)
			for (int n = 0; n < Nc; ++n) tw1[m * Nc + n] = cv1[n * Np + m] * c->w_vc[n];
		for (int k = 0; k < d; ++k) ops[k] = (k == dim) ? Dg : Id;
		or_kron(d, ops, rowsN, rowsN, c->d_vm + (size_t)dim * Np * Np);
	}

	c->w_fc = (double*)malloc(sizeof(double) * Nfc);
	for (int n = 0; n < Nfc; ++n) {
		int r = n;
		double w = 1.0;
		for (int k = 0; k < d - 1; ++k) { w *= c->wgl[r % N]; r /= N; }
		c->w_fc[n] = w;
	}
	c->cv0_vs_fc = (double*)malloc(sizeof(double) * 2 * d * Nfc * Np);
	c->tw0_vt_fc = (double*)malloc(sizeof(double) * 2 * d * Np * Nfc);
	for (int lf = 0; lf < 2 * d; ++lf) {
		const int fd = lf / 2;
		int rows[3], cols[3];
		for (int k = 0; k < d; ++k) {
			ops[k] = (k == fd) ? ((lf % 2) ? oneN : one0) : I1;
			rows[k] = (k == fd) ? 1 : N;
			cols[k] = N;
		}
		double* cv = c->cv0_vs_fc + (size_t)lf * Nfc * Np;
		or_kron(d, ops, rows, cols, cv);
		double* tw = c->tw0_vt_fc + (size_t)lf * Np * Nfc;
		for (int m = 0; m < Np; ++m)
			for (int n = 0; n < Nfc; ++n) tw[m * Nfc + n] = cv[n * Np + m] * c->w_fc[n];
	}
	(void)face_vol_node;

	c->n_ord = (d == 3) ? 8 : (d == 2 ? 2 : 1);
	c->perm = (int*)malloc(sizeof(int) * c->n_ord * Nfc);
	for (int o = 0; o < c->n_ord; ++o) or_face_perm(c, o, c->perm + (size_t)o * Nfc);

	c->faces = (int*)malloc(sizeof(int) * 6 * (F > 0 ? F : 1));
	memcpy(c->faces, faces, sizeof(int) * 6 * F);
	c->xyz = (double*)malloc(sizeof(double) * (size_t)E * d * Np);
	memcpy(c->xyz, xyz, sizeof(double) * (size_t)E * d * Np);
	c->metrics_vm = (double*)malloc(sizeof(double) * (size_t)E * d * d * Np);
	c->metrics_vc = (double*)malloc(sizeof(double) * (size_t)E * d * d * Nc);
	c->detJ_vc = (double*)malloc(sizeof(double) * (size_t)E * Nc);
	c->nJ_fc = (double*)malloc(sizeof(double) * (size_t)(F > 0 ? F : 1) * Nfc * d);
	compute_geometry(c);

	/* per-element face lists */
	c->ef_off = (int*)calloc(E + 1, sizeof(int));
	for (int f = 0; f < F; ++f) {
		c->ef_off[faces[6 * f] + 1]++;
		if (faces[6 * f + 5] == OR_BC_INTERIOR) c->ef_off[faces[6 * f + 2] + 1]++;
	}
	for (int e = 0; e < E; ++e) c->ef_off[e + 1] += c->ef_off[e];
	c->ef_list = (int*)malloc(sizeof(int) * (c->ef_off[E] > 0 ? c->ef_off[E] : 1));
	int* fill = (int*)calloc(E, sizeof(int));
	for (int f = 0; f < F; ++f) {
		const int eL = faces[6 * f];
		c->ef_list[c->ef_off[eL] + fill[eL]++] = (f << 1) | 0;
		if (faces[6 * f + 5] == OR_BC_INTERIOR) {
			const int eR = faces[6 * f + 2];
			c->ef_list[c->ef_off[eR] + fill[eR]++] = (f << 1) | 1;
		}
	}
	free(fill);

	/* Bezier <-> nodal (ccBS0 / ccSB0, used by the positivity limiter) */
	if (d <= 3) {
		double B1[256];
		or_bernstein(p, N, c->xgll, B1);
		for (int k = 0; k < d; ++k) ops[k] = B1;
		c->ccBS = (double*)malloc(sizeof(double) * Np * Np);
		c->ccSB = (double*)malloc(sizeof(double) * Np * Np);
		or_kron(d, ops, rowsN, rowsN, c->ccBS);
		invert(Np, c->ccBS, c->ccSB);
	}
	return c;
}

void or_destroy(or_ctx* c)
{
	if (!c) return;
	free(c->faces); free(c->cv0_vs_vc); free(c->cv1_vs_vc); free(c->tw1_vt_vc); free(c->cv0_vs_fc);
	free(c->tw0_vt_fc); free(c->d_vm); free(c->w_vc); free(c->w_fc); free(c->perm); free(c->xyz);
	free(c->metrics_vm); free(c->metrics_vc); free(c->detJ_vc); free(c->nJ_fc); free(c->minv);
	free(c->ccBS); free(c->ccSB); free(c->ef_off); free(c->ef_list);
	free(c);
}

int or_np(const or_ctx* c) { return c->Np; }
int or_nv(const or_ctx* c) { return c->nv; }

void or_geometry(const or_ctx* c, double* metrics_vc, double* detJ_vc, double* nJ_fc)
{
	const int d = c->d;
	if (metrics_vc) memcpy(metrics_vc, c->metrics_vc, sizeof(double) * (size_t)c->E * d * d * c->Nc);
	if (detJ_vc) memcpy(detJ_vc, c->detJ_vc, sizeof(double) * (size_t)c->E * c->Nc);
	if (nJ_fc) memcpy(nJ_fc, c->nJ_fc, sizeof(double) * (size_t)c->F * c->Nfc * d);
}

void or_compute_rhs_d(const or_ctx* c, const double* u, double* rhs)
{
	memset(rhs, 0, sizeof(double) * (size_t)c->E * c->Np * c->nv);
	volume_rhs_d(c, u, rhs);
	face_rhs_d(c, u, rhs);
}

/* The same residual on nthreads OpenMP threads (the CPU baseline's all-cores figure; the reference itself is
 * single-threaded).  Volume terms are element-local.  The face fluxes are computed in parallel into a per-face
 * buffer, then lifted colour by colour: faces of one colour (greedy over the face list) share no element, so
 * their lifts touch disjoint residual blocks.  Each element receives the same additions in the same order as
 * or_compute_rhs_d up to the face order, so the result agrees to rounding. */
/* greedy face colouring: faces of one colour share no element (their lifts touch disjoint residual blocks) */
static int face_colors(const or_ctx* c, int* color)
{
	unsigned long long* used = (unsigned long long*)calloc((size_t)c->E, sizeof(unsigned long long));
	int ncol = 0;
	for (int f = 0; f < c->F; ++f) {
		const int* fc = c->faces + 6 * f;
		unsigned long long m = used[fc[0]] | (fc[2] >= 0 ? used[fc[2]] : 0ull);
		int k = 0;
		while (k < 63 && ((m >> k) & 1ull)) ++k;
		color[f] = k;
		used[fc[0]] |= 1ull << k;
		if (fc[2] >= 0) used[fc[2]] |= 1ull << k;
		if (k + 1 > ncol) ncol = k + 1;
	}
	free(used);
	return ncol;
}

void or_compute_rhs_d_threads(const or_ctx* c, const double* u, double* rhs, int nthreads)
{
	const size_t blk = (size_t)c->Np * c->nv, fsz = (size_t)c->Nfc * c->nv;
	if (nthreads < 1) nthreads = 1;
	memset(rhs, 0, sizeof(double) * (size_t)c->E * blk);
	double* nnf = (double*)malloc(sizeof(double) * (size_t)c->F * fsz);
	int* color = (int*)malloc(sizeof(int) * (size_t)c->F);
	const int ncol = face_colors(c, color);
#pragma omp parallel num_threads(nthreads)
	{
#pragma omp for schedule(static)
		for (int e = 0; e < c->E; ++e) volume_elem_d(c, e, u, rhs);
#pragma omp for schedule(static)
		for (int f = 0; f < c->F; ++f) face_nnf_d(c, f, u, nnf + (size_t)f * fsz);
		for (int k = 0; k < ncol; ++k) {
#pragma omp for schedule(static)
			for (int f = 0; f < c->F; ++f) {
				if (color[f] != k) continue;
				face_side_d(c, f, 0, nnf + (size_t)f * fsz, rhs);
				if (c->faces[6 * f + 5] == OR_BC_INTERIOR) face_side_d(c, f, 1, nnf + (size_t)f * fsz, rhs);
			}
		}
	}
	free(color);
	free(nnf);
}

void or_volume_rhs_d(const or_ctx* c, const double* u, double* rhs)
{
	memset(rhs, 0, sizeof(double) * (size_t)c->E * c->Np * c->nv);
	volume_rhs_d(c, u, rhs);
}

void or_face_rhs_d(const or_ctx* c, const double* u, double* rhs)
{
	memset(rhs, 0, sizeof(double) * (size_t)c->E * c->Np * c->nv);
	face_rhs_d(c, u, rhs);
}

void or_compute_rhs_c(const or_ctx* c, const double* ure, const double* uim, double* rre, double* rim)
{
	const size_t n = (size_t)c->E * c->Np * c->nv;
	double complex* u = (double complex*)malloc(sizeof(double complex) * n);
	double complex* r = (double complex*)calloc(n, sizeof(double complex));
	for (size_t i = 0; i < n; ++i) u[i] = ure[i] + I * (uim ? uim[i] : 0.0);
	volume_rhs_c(c, u, r);
	face_rhs_c(c, u, r);
	for (size_t i = 0; i < n; ++i) { rre[i] = creal(r[i]); rim[i] = cimag(r[i]); }
	free(u); free(r);
}

/* The complex twin on nthreads threads (the CPU Newton-GMRES's complex-step J.v), organised as
 * or_compute_rhs_d_threads. */
void or_compute_rhs_c_threads(const or_ctx* c, const double* ure, const double* uim, double* rre, double* rim,
                              int nthreads)
{
	const size_t n = (size_t)c->E * c->Np * c->nv, fsz = (size_t)c->Nfc * c->nv;
	if (nthreads < 1) nthreads = 1;
	double complex* u = (double complex*)malloc(sizeof(double complex) * n);
	double complex* r = (double complex*)calloc(n, sizeof(double complex));
	double complex* nnf = (double complex*)malloc(sizeof(double complex) * (size_t)c->F * fsz);
	int* color = (int*)malloc(sizeof(int) * (size_t)c->F);
	const int ncol = face_colors(c, color);
#pragma omp parallel num_threads(nthreads)
	{
#pragma omp for schedule(static)
		for (size_t i = 0; i < n; ++i) u[i] = ure[i] + I * (uim ? uim[i] : 0.0);
#pragma omp for schedule(static)
		for (int e = 0; e < c->E; ++e) volume_elem_c(c, e, u, r);
#pragma omp for schedule(static)
		for (int f = 0; f < c->F; ++f) face_nnf_c(c, f, u, nnf + (size_t)f * fsz);
		for (int k = 0; k < ncol; ++k) {
#pragma omp for schedule(static)
			for (int f = 0; f < c->F; ++f) {
				if (color[f] != k) continue;
				face_side_c(c, f, 0, nnf + (size_t)f * fsz, r);
				if (c->faces[6 * f + 5] == OR_BC_INTERIOR) face_side_c(c, f, 1, nnf + (size_t)f * fsz, r);
			}
		}
#pragma omp for schedule(static)
		for (size_t i = 0; i < n; ++i) { rre[i] = creal(r[i]); rim[i] = cimag(r[i]); }
	}
	free(u); free(r); free(nnf); free(color);
}

/* mass matrix cv0^T diag(w detJ) cv0 -- volume_solver_T.c:115-131 */
void or_mass(const or_ctx* c, int e, double* M)
{
	const int Np = c->Np, Nc = c->Nc;
	const double* dj = c->detJ_vc + (size_t)e * Nc;
	for (int i = 0; i < Np; ++i)
		for (int j = 0; j < Np; ++j) {
			double acc = 0.0;
			for (int n = 0; n < Nc; ++n) acc += c->cv0_vs_vc[n * Np + i] * (c->w_vc[n] * dj[n]) * c->cv0_vs_vc[n * Np + j];
			M[i * Np + j] = acc;
		}
}

int or_apply_minv(or_ctx* c, double* rhs)
{
	const int Np = c->Np, nv = c->nv;
	if (!c->minv) {
		c->minv = (double*)malloc(sizeof(double) * (size_t)c->E * Np * Np);
		double* M = (double*)malloc(sizeof(double) * Np * Np);
		for (int e = 0; e < c->E; ++e) {
			or_mass(c, e, M);
			if (invert(Np, M, c->minv + (size_t)e * Np * Np)) { free(M); return -1; }
		}
		free(M);
	}
	double* tmp = (double*)malloc(sizeof(double) * Np);
	for (int e = 0; e < c->E; ++e) {
		const double* Mi = c->minv + (size_t)e * Np * Np;
		for (int v = 0; v < nv; ++v) {
			double* r = rhs + ((size_t)e * nv + v) * Np;
			for (int i = 0; i < Np; ++i) {
				double acc = 0.0;
				for (int j = 0; j < Np; ++j) acc += Mi[i * Np + j] * r[j];
				tmp[i] = acc;
			}
			memcpy(r, tmp, sizeof(double) * Np);
		}
	}
	free(tmp);
	return 0;
}

/* ------------------------------------------------------------------------------------------------ */
/* Navier-Stokes with BR2 (real only).                                                              */

void or_set_viscous(or_ctx* c, double mu, double Pr)
{
	c->viscous = 1;
	c->mu = mu;
	c->Pr = Pr;
}

void or_set_walls(or_ctx* c, double nf_E, double omega, double rho_b, double E_b)
{
	c->wall_nfE = nf_E;
	c->wall_omega = omega;
	c->wall_rho = rho_b;
	c->wall_E = E_b;
}

void or_set_sutherland(or_ctx* c, double r_s)
{
	c->sutherland = 1;
	c->r_s = r_s;
}

static int ns_bc_supported(int bc)
{
	return bc == OR_BC_RIEMANN || bc == OR_BC_NOSLIP_ADIABATIC || bc == OR_BC_NOSLIP_DIABATIC || bc == OR_BC_NOSLIP_ROTATING;
}

/* physical coordinates at the face cubature nodes of local face lf of element e (isoparametric): xy[n * d + k] */
static void face_xyz(const or_ctx* c, int e, int lf, double* xy)
{
	const int d = c->d, Np = c->Np, Nfc = c->Nfc;
	const double* op = c->cv0_vs_fc + (size_t)lf * Nfc * Np;
	for (int k = 0; k < d; ++k)
		for (int n = 0; n < Nfc; ++n) {
			double acc = 0.0;
			for (int m = 0; m < Np; ++m) acc += op[n * Np + m] * c->xyz[((size_t)e * d + k) * Np + m];
			xy[n * d + k] = acc;
		}
}

static int ensure_minv(or_ctx* c)
{
	const int Np = c->Np;
	if (c->minv) return 0;
	c->minv = (double*)malloc(sizeof(double) * (size_t)c->E * Np * Np);
	double* M = (double*)malloc(sizeof(double) * Np * Np);
	for (int e = 0; e < c->E; ++e) {
		or_mass(c, e, M);
		if (invert(Np, M, c->minv + (size_t)e * Np * Np)) { free(M); free(c->minv); c->minv = NULL; return -1; }
	}
	free(M);
	return 0;
}

void or_viscous_flux(int d, const double* s, const double* g, double mu, double Pr, double* fv)
{
	viscous_flux_d(d, s, g, mu, Pr, fv);
}

int or_ns_gradient(or_ctx* c, const double* u, double* g, double* gv, double* gf)
{
	return ns_gradient_d(c, u, g, gv, gf);
}

int or_compute_rhs_ns(or_ctx* c, const double* u, double* rhs)
{
	return compute_rhs_ns_d(c, u, rhs);
}

/* the complex twin (complex-step derivatives of the BR2 residual: J v = Im R(u + i h v) / h) */
int or_compute_rhs_ns_c(or_ctx* c, const double* ure, const double* uim, double* rre, double* rim)
{
	const size_t n = (size_t)c->E * c->Np * c->nv;
	double complex* u = (double complex*)calloc(n ? n : 1, sizeof(double complex));
	double complex* r = (double complex*)malloc(sizeof(double complex) * (n ? n : 1));
	for (size_t i = 0; i < n; ++i) u[i] = ure[i] + I * (uim ? uim[i] : 0.0);
	const int st = compute_rhs_ns_c(c, u, r);
	if (!st)
		for (size_t i = 0; i < n; ++i) { rre[i] = creal(r[i]); rim[i] = cimag(r[i]); }
	free(u); free(r);
	return st;
}

double or_compute_rhs_explicit(or_ctx* c, const double* u, double* rhs)
{
	or_compute_rhs_d(c, u, rhs);
	or_apply_minv(c, rhs);
	const size_t n = (size_t)c->E * c->Np * c->nv;
	double mx = 0.0;
	for (size_t i = 0; i < n; ++i) if (fabs(rhs[i]) > mx) mx = fabs(rhs[i]);
	return mx;
}

/* enforce_positivity_highorder -- solve.c:187-233, correct_coef solve.c:356-374 */
int or_enforce_positivity(const or_ctx* c, double* u)
{
	if (c->pde != OR_PDE_EULER) return 0;
	const int d = c->d, Np = c->Np, nv = c->nv;
	double* b = (double*)malloc(sizeof(double) * Np * nv);
	int status = 0;
	for (int e = 0; e < c->E; ++e) {
		double* ue = u + (size_t)e * Np * nv;
		for (int v = 0; v < nv; ++v)
			for (int i = 0; i < Np; ++i) {
				double acc = 0.0;
				for (int j = 0; j < Np; ++j) acc += c->ccSB[i * Np + j] * ue[v * Np + j];
				b[v * Np + i] = acc;
			}
		/* conserved -> primitive on the Bezier coefficients */
		for (int n = 0; n < Np; ++n) {
			const double rho = b[n];
			double V2 = 0.0;
			for (int k = 0; k < d; ++k) { b[(1 + k) * Np + n] /= rho; V2 += b[(1 + k) * Np + n] * b[(1 + k) * Np + n]; }
			b[(d + 1) * Np + n] = OR_GM1 * (b[(d + 1) * Np + n] - 0.5 * rho * V2);
		}
		for (int vr = 0; vr < nv; vr += nv - 1) {
			const double* vd = &b[vr * Np];
			double avg = 0.0, mn = vd[0];
			for (int n = 0; n < Np; ++n) { avg += vd[n]; if (vd[n] < mn) mn = vd[n]; }
			avg /= Np;
			if (avg < OR_EPS_PHYS) { status = -1; continue; }
			if (mn < OR_EPS_PHYS) {
				double p_ho = (avg - OR_EPS_PHYS) / (avg - mn);
				p_ho = p_ho < 0.0 ? 0.0 : (p_ho > 1.0 ? 1.0 : p_ho);
				for (int v2 = 0; v2 < nv; ++v2) {
					double* dd = &b[v2 * Np];
					double a2 = 0.0;
					for (int n = 0; n < Np; ++n) a2 += dd[n];
					a2 /= Np;
					if ((v2 == 0 || v2 == nv - 1) && a2 < OR_EPS_PHYS) a2 = OR_EPS_PHYS;
					for (int n = 0; n < Np; ++n) dd[n] = p_ho * dd[n] + (1 - p_ho) * a2;
				}
			}
		}
		/* primitive -> conserved */
		for (int n = 0; n < Np; ++n) {
			const double rho = b[n];
			double V2 = 0.0;
			for (int k = 0; k < d; ++k) V2 += b[(1 + k) * Np + n] * b[(1 + k) * Np + n];
			for (int k = 0; k < d; ++k) b[(1 + k) * Np + n] *= rho;
			b[(d + 1) * Np + n] = b[(d + 1) * Np + n] / OR_GM1 + 0.5 * rho * V2;
		}
		for (int v = 0; v < nv; ++v)
			for (int i = 0; i < Np; ++i) {
				double acc = 0.0;
				for (int j = 0; j < Np; ++j) acc += c->ccBS[i * Np + j] * b[v * Np + j];
				ue[v * Np + i] = acc;
			}
	}
	free(b);
	return status;
}

/* time_step_ssp_rk_33 -- solve_explicit.c:252-301 */
double or_ssp_rk3_step(or_ctx* c, double* u, double dt, int positivity, int* err)
{
	const size_t n = (size_t)c->E * c->Np * c->nv;
	double* up = (double*)malloc(sizeof(double) * n);
	double* r = (double*)malloc(sizeof(double) * n);
	double max_rhs = 0.0;
	if (err) *err = 0;
	for (int rk = 0; rk < 3; ++rk) {
		max_rhs = or_compute_rhs_explicit(c, u, r);
		if (rk == 0) {
			for (size_t i = 0; i < n; ++i) { up[i] = u[i]; u[i] += dt * r[i]; }
		} else if (rk == 1) {
			for (size_t i = 0; i < n; ++i) u[i] = (1.0 / 4.0) * (3.0 * up[i] + u[i] + dt * r[i]);
		} else {
			for (size_t i = 0; i < n; ++i) u[i] = (1.0 / 3.0) * (up[i] + 2.0 * u[i] + 2.0 * dt * r[i]);
		}
		if (positivity && or_enforce_positivity(c, u) && err) *err = -1;
	}
	free(up); free(r);
	return max_rhs;
}

void or_euler_flux_d(int d, const double* s, double* f) { euler_flux_d(d, s, f); }

void or_numflux_d(int d, int pde, int flux, const double* sL, const double* sR, const double* n,
                  const double* b_adv, double* nnf)
{
	numflux_d(d, pde, flux, sL, sR, n, b_adv, nnf);
}

void or_numflux_c(int d, int pde, int flux, const double* sLre, const double* sLim, const double* sRre,
                  const double* sRim, const double* n, const double* b_adv, double* nre, double* nim)
{
	const int nv = (pde == OR_PDE_EULER) ? d + 2 : 1;
	double complex sL[5], sR[5], out[5];
	for (int v = 0; v < nv; ++v) { sL[v] = sLre[v] + I * sLim[v]; sR[v] = sRre[v] + I * sRim[v]; }
	numflux_c(d, pde, flux, sL, sR, n, b_adv, out);
	for (int v = 0; v < nv; ++v) { nre[v] = creal(out[v]); nim[v] = cimag(out[v]); }
}

void or_bc_state_d(int d, int bc, const double* sL, const double* n, const double* bc_prim, double* sR)
{
	bc_state_d(d, bc, sL, n, bc_prim, sR);
}

/* complex-step Jacobian block -- test_support_solve_dg.c:164-212 (column by column, s += i*h) */
void or_jacobian_block_cs(const or_ctx* c, const double* u, int e_row, int e_col, double h, double* blk)
{
	const size_t ntot = (size_t)c->E * c->Np * c->nv;
	const int nb = c->Np * c->nv;
	double complex* uc = (double complex*)malloc(sizeof(double complex) * ntot);
	double complex* rc = (double complex*)calloc(ntot, sizeof(double complex));
	for (size_t i = 0; i < ntot; ++i) uc[i] = u[i];
	for (int j = 0; j < nb; ++j) {
		const size_t gj = (size_t)e_col * nb + j;
		uc[gj] = u[gj] + I * h;
		rhs_local_c(c, e_row, uc, rc);
		uc[gj] = u[gj];
		for (int i = 0; i < nb; ++i) blk[(size_t)i * nb + j] = cimag(rc[(size_t)e_row * nb + i]) / h;
	}
	free(uc); free(rc);
}

void or_diag_blocks_cs(const or_ctx* c, const double* u, double h, double* blocks)
{
	const int nb = c->Np * c->nv;
	for (int e = 0; e < c->E; ++e) or_jacobian_block_cs(c, u, e, e, h, blocks + (size_t)e * nb * nb);
}

/* The same on nthreads threads, each with its own complex copy of the state (a perturbed column is visible to
 * the faces of the neighbours, so the copies cannot be shared). */
void or_diag_blocks_cs_threads(const or_ctx* c, const double* u, double h, double* blocks, int nthreads)
{
	const size_t ntot = (size_t)c->E * c->Np * c->nv;
	const int nb = c->Np * c->nv;
	if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
	{
		double complex* uc = (double complex*)malloc(sizeof(double complex) * ntot);
		double complex* rc = (double complex*)calloc(ntot, sizeof(double complex));
		for (size_t i = 0; i < ntot; ++i) uc[i] = u[i];
#pragma omp for schedule(dynamic, 16)
		for (int e = 0; e < c->E; ++e) {
			double* blk = blocks + (size_t)e * nb * nb;
			for (int j = 0; j < nb; ++j) {
				const size_t gj = (size_t)e * nb + j;
				uc[gj] = u[gj] + I * h;
				rhs_local_c(c, e, uc, rc);
				uc[gj] = u[gj];
				for (int i = 0; i < nb; ++i) blk[(size_t)i * nb + j] = cimag(rc[(size_t)e * nb + i]) / h;
			}
		}
		free(uc); free(rc);
	}
}

/* Bezier coefficients of the nv variables of element e (ccSB0): b[v * Np + i] */
static void bezier_coefs(const or_ctx* c, const double* ue, double* b)
{
	const int Np = c->Np, nv = c->nv;
	for (int v = 0; v < nv; ++v)
		for (int i = 0; i < Np; ++i) {
			double acc = 0.0;
			for (int j = 0; j < Np; ++j) acc += c->ccSB[i * Np + j] * ue[v * Np + j];
			b[v * Np + i] = acc;
		}
}

/* 1 / dt per element for the pseudo-time step: dt = cfl * h_min / max(|V| + c) over the Bezier coefficients of
 * the state (compute_dt_cfl_constrained, solve_dg.c:284-319), h_min the shortest straight edge between the
 * element's vertices (compute_min_length_measure). */
void or_cfl_inv_dt(const or_ctx* c, const double* u, double cfl, double* inv_dt)
{
	const int d = c->d, Np = c->Np, nv = c->nv, N = c->N;
	double* b = (double*)malloc(sizeof(double) * Np * nv);
	for (int e = 0; e < c->E; ++e) {
		bezier_coefs(c, u + (size_t)e * Np * nv, b);
		double m = 0.0;
		for (int i = 0; i < Np; ++i) {
			const double rho = b[i], ri = 1.0 / rho;
			double V2 = 0.0;
			for (int k = 0; k < d; ++k) { const double w = b[(1 + k) * Np + i] * ri; V2 += w * w; }
			const double p = OR_GM1 * (b[(d + 1) * Np + i] - 0.5 * rho * V2);
			const double ws = sqrt(V2) + sqrt(OR_GAMMA * p * ri);
			if (ws > m) m = ws;
		}
		double hmin = 1e300;
		for (int a = 0; a < (1 << d); ++a)
			for (int k = 0; k < d; ++k) {
				const int bb = a ^ (1 << k);
				if (bb < a) continue;
				int na = 0, nb2 = 0, st = 1;
				for (int q = 0; q < d; ++q) {
					na += ((a >> q) & 1) * (N - 1) * st;
					nb2 += ((bb >> q) & 1) * (N - 1) * st;
					st *= N;
				}
				double s2 = 0.0;
				for (int x = 0; x < d; ++x) {
					const double t = c->xyz[((size_t)e * d + x) * Np + na] - c->xyz[((size_t)e * d + x) * Np + nb2];
					s2 += t * t;
				}
				if (sqrt(s2) < hmin) hmin = sqrt(s2);
			}
		inv_dt[e] = m / (cfl * hmin);
		/* Navier-Stokes: the viscous bound dt <= cfl h^2 / mu_max, mu_max over the Bezier coefficients
		 * (solve_dg.c:300-318) */
		if (c->viscous) {
			double mmu = c->mu;
			if (c->sutherland) {
				mmu = 0.0;
				for (int i = 0; i < Np; ++i) {
					double s[5];
					for (int v = 0; v < nv; ++v) s[v] = b[v * Np + i];
					const double m = mu_of_d(c, s);
					if (m > mmu) mmu = m;
				}
			}
			if (mmu / (cfl * hmin * hmin) > inv_dt[e]) inv_dt[e] = mmu / (cfl * hmin * hmin);
		}
	}
	free(b);
}

/* u_e += relax_e du_e with relax halved until the means of the Bezier coefficients of rho and p exceed EPS_PHYS
 * (compute_relaxation, solve_implicit.c:823-873).  Returns 0, or -1 when an element needs relax < EPS. */
int or_relaxed_update(const or_ctx* c, double* u, const double* du)
{
	const int d = c->d, Np = c->Np, nv = c->nv;
	double* bu = (double*)malloc(sizeof(double) * Np * nv);
	double* bd = (double*)malloc(sizeof(double) * Np * nv);
	int status = 0;
	for (int e = 0; e < c->E; ++e) {
		double* ue = u + (size_t)e * Np * nv;
		const double* de = du + (size_t)e * Np * nv;
		bezier_coefs(c, ue, bu);
		bezier_coefs(c, de, bd);
		double relax = 1.0;
		for (;;) {
			double ar = 0.0, ap = 0.0;
			for (int i = 0; i < Np; ++i) {
				const double rho = bu[i] + relax * bd[i];
				double V2 = 0.0;
				for (int k = 0; k < d; ++k) {
					const double w = (bu[(1 + k) * Np + i] + relax * bd[(1 + k) * Np + i]) / rho;
					V2 += w * w;
				}
				ar += rho;
				ap += OR_GM1 * ((bu[(d + 1) * Np + i] + relax * bd[(d + 1) * Np + i]) - 0.5 * rho * V2);
			}
			ar /= Np;
			ap /= Np;
			if (!(ar < OR_EPS_PHYS || ap < OR_EPS_PHYS)) break;
			relax *= 0.5;
			if (relax < OR_EPS) { status = -1; break; }
		}
		for (int i = 0; i < Np * nv; ++i) ue[i] += relax * de[i];
	}
	free(bu); free(bd);
	return status;
}

/* Ainv_e = A_e^-1 for E dense nb x nb blocks (row-major; Gauss-Jordan with partial pivoting), on nthreads
 * threads.  Returns -1 if a block is singular. */
int or_block_inverse(int E, int nb, const double* A, double* Ainv, int nthreads)
{
	int bad = 0;
	if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
	{
		double* a = (double*)malloc(sizeof(double) * (size_t)nb * nb);
#pragma omp for schedule(static) reduction(|| : bad)
		for (int e = 0; e < E; ++e) {
			memcpy(a, A + (size_t)e * nb * nb, sizeof(double) * (size_t)nb * nb);
			double* x = Ainv + (size_t)e * nb * nb;
			if (invert(nb, a, x)) bad = 1;
		}
		free(a);
	}
	return bad ? -1 : 0;
}

/* h_j = V_j . w for the first k rows of V (k x n, row-major), on nthreads threads */
void or_dots(int k, long long n, const double* V, const double* w, double* h, int nthreads)
{
	if (nthreads < 1) nthreads = 1;
	for (int j = 0; j < k; ++j) h[j] = 0.0;
#pragma omp parallel num_threads(nthreads)
	{
		double loc[256];
		for (int j = 0; j < k; ++j) loc[j] = 0.0;
#pragma omp for schedule(static)
		for (long long i = 0; i < n; ++i)
			for (int j = 0; j < k; ++j) loc[j] += V[(size_t)j * n + i] * w[i];
#pragma omp critical
		for (int j = 0; j < k; ++j) h[j] += loc[j];
	}
}

/* w -= sum_j h_j V_j over the first k rows of V, on nthreads threads */
void or_gemv_sub(int k, long long n, const double* V, const double* h, double* w, int nthreads)
{
	if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
	for (long long i = 0; i < n; ++i) {
		double acc = w[i];
		for (int j = 0; j < k; ++j) acc -= h[j] * V[(size_t)j * n + i];
		w[i] = acc;
	}
}

/* z_e = A_e v_e for E dense nb x nb blocks (row-major), on nthreads threads */
void or_block_apply(int E, int nb, const double* A, const double* v, double* z, int nthreads)
{
	if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
	for (int e = 0; e < E; ++e) {
		const double* Ae = A + (size_t)e * nb * nb;
		const double* ve = v + (size_t)e * nb;
		double* ze = z + (size_t)e * nb;
		for (int i = 0; i < nb; ++i) {
			double acc = 0.0;
			for (int j = 0; j < nb; ++j) acc += Ae[(size_t)i * nb + j] * ve[j];
			ze[i] = acc;
		}
	}
}
