// dgx_tri.cpp -- host construction of the TRI element of the reference's mixed 2-D meshes.
//
// Reference element (src/element/nodes.c:628-852): the equilateral triangle with vertices V0 = (-1, -1/sqrt 3),
// V1 = (1, -1/sqrt 3), V2 = (0, 2/sqrt 3); symmetric node sets given per orbit in barycentric coordinates,
// expanded orbit by orbit through the permutations (0,1,2) (2,0,1) (1,2,0) (0,2,1) (1,0,2) (2,1,0).
// Solution nodes: WSH of order p (interp_si WSH), Lagrange basis; volume cubature: WV of degree 2p
// (nodes_operators.c:327-345, 436-520); face cubature: GL of the face order.  Local face lf is opposite vertex
// lf and runs from f_ve[lf][0] to f_ve[lf][1] = (1,2), (0,2), (0,1) (element.c:454-463).
//
// The Lagrange operators are formed through monomials r^a s^b with the Vandermonde inverse in long double, so
// that the double operators are rounded once (Np <= 21).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <vector>

#include "dgx_internal.h"

namespace {

#include "dgx_tri_tables.inc"

const double kSqrt3 = 1.7320508075688772935;
const double kTriV[3][2] = { { -1.0, -1.0 / kSqrt3 }, { 1.0, -1.0 / kSqrt3 }, { 0.0, 2.0 / kSqrt3 } };
const int kPerm[6][3] = { { 0, 1, 2 }, { 2, 0, 1 }, { 1, 2, 0 }, { 0, 2, 1 }, { 1, 0, 2 }, { 2, 1, 0 } };
const int kFaceVe[3][2] = { { 1, 2 }, { 0, 2 }, { 0, 1 } };

const TriRule* find_rule(const char* kind, int q)
{
	char name[16];
	std::snprintf(name, sizeof name, "%s%d", kind, q);
	for (const TriRule& t : kTriRules)
		if (std::strcmp(t.name, name) == 0) return &t;
	return nullptr;
}

// orbits -> points, in the reference's order; rst = bary . V
int expand(const TriRule& t, std::vector<double>& rs, std::vector<double>& w, std::vector<double>* bary = nullptr)
{
	int ig = 0;
	while (ig < 3 && t.symms[ig][0] == 0) ++ig;
	int count = 0;
	rs.clear();
	w.clear();
	if (bary) bary->clear();
	for (int g = 0; g < t.ngroups; ++g) {
		if (ig >= 3) return -1;
		++count;
		for (int j = 0; j < t.symms[ig][1]; ++j) {
			double b[3];
			for (int k = 0; k < 3; ++k) b[k] = t.bary[g][kPerm[j][k]];
			rs.push_back(b[0] * kTriV[0][0] + b[1] * kTriV[1][0] + b[2] * kTriV[2][0]);
			rs.push_back(b[0] * kTriV[0][1] + b[1] * kTriV[1][1] + b[2] * kTriV[2][1]);
			w.push_back(t.w[g]);
			if (bary) bary->insert(bary->end(), b, b + 3);
		}
		if (count == t.symms[ig][0]) {
			count = 0;
			++ig;
		}
	}
	return (int)w.size();
}

// monomials r^a s^b, a + b <= p (total degree, then b), with their r and s derivatives
void monomials(int p, double r, double s, long double* v, long double* vr, long double* vs)
{
	int k = 0;
	for (int tot = 0; tot <= p; ++tot)
		for (int b = 0; b <= tot; ++b, ++k) {
			const int a = tot - b;
			const long double ra = std::pow((long double)r, a), sb = std::pow((long double)s, b);
			v[k] = ra * sb;
			vr[k] = a ? a * std::pow((long double)r, a - 1) * sb : 0.0L;
			vs[k] = b ? b * ra * std::pow((long double)s, b - 1) : 0.0L;
		}
}

bool invert_ld(int n, std::vector<long double>& A, std::vector<long double>& Ainv)
{
	std::vector<long double> M((size_t)n * 2 * n);
	for (int i = 0; i < n; ++i)
		for (int j = 0; j < 2 * n; ++j) M[(size_t)i * 2 * n + j] = j < n ? A[(size_t)i * n + j] : (j - n == i ? 1.0L : 0.0L);
	for (int c = 0; c < n; ++c) {
		int piv = c;
		for (int i = c + 1; i < n; ++i)
			if (std::fabs(M[(size_t)i * 2 * n + c]) > std::fabs(M[(size_t)piv * 2 * n + c])) piv = i;
		if (M[(size_t)piv * 2 * n + c] == 0.0L) return false;
		if (piv != c)
			for (int j = 0; j < 2 * n; ++j) std::swap(M[(size_t)c * 2 * n + j], M[(size_t)piv * 2 * n + j]);
		const long double d = 1.0L / M[(size_t)c * 2 * n + c];
		for (int j = 0; j < 2 * n; ++j) M[(size_t)c * 2 * n + j] *= d;
		for (int i = 0; i < n; ++i) {
			if (i == c) continue;
			const long double f = M[(size_t)i * 2 * n + c];
			if (f == 0.0L) continue;
			for (int j = 0; j < 2 * n; ++j) M[(size_t)i * 2 * n + j] -= f * M[(size_t)c * 2 * n + j];
		}
	}
	Ainv.assign((size_t)n * n, 0.0L);
	for (int i = 0; i < n; ++i)
		for (int j = 0; j < n; ++j) Ainv[(size_t)i * n + j] = M[(size_t)i * 2 * n + n + j];
	return true;
}

// the order-p Lagrange basis on the WSH nodes: Vinv (Np x Np, monomial coefficients of each basis function)
struct TriBasis {
	int p = 0, Np = 0;
	std::vector<double> rs, w;
	std::vector<long double> Vinv;
	bool init(int p_)
	{
		p = p_;
		const TriRule* t = find_rule("WSH", p);
		if (!t) return false;
		Np = expand(*t, rs, w);
		if (Np != (p + 1) * (p + 2) / 2) return false;
		std::vector<long double> V((size_t)Np * Np), vr(Np), vs(Np);
		for (int i = 0; i < Np; ++i) monomials(p, rs[2 * i], rs[2 * i + 1], &V[(size_t)i * Np], vr.data(), vs.data());
		return invert_ld(Np, V, Vinv);
	}
	// phi_m, d phi_m / dr, d phi_m / ds at (r, s)
	void eval(double r, double s, double* phi, double* dr, double* ds) const
	{
		std::vector<long double> v(Np), vr(Np), vs(Np);
		monomials(p, r, s, v.data(), vr.data(), vs.data());
		for (int m = 0; m < Np; ++m) {
			long double a = 0.0L, b = 0.0L, c = 0.0L;
			for (int k = 0; k < Np; ++k) {
				const long double ck = Vinv[(size_t)k * Np + m];   // V phi-coefficients: phi_m = sum_k Vinv[k][m] x^k
				a += v[k] * ck;
				b += vr[k] * ck;
				c += vs[k] * ck;
			}
			if (phi) phi[m] = (double)a;
			if (dr) dr[m] = (double)b;
			if (ds) ds[m] = (double)c;
		}
	}
};

void face_point(int lf, double s, double* rs, double* drs)
{
	const double* a = kTriV[kFaceVe[lf][0]];
	const double* b = kTriV[kFaceVe[lf][1]];
	rs[0] = 0.5 * (1.0 - s) * a[0] + 0.5 * (1.0 + s) * b[0];
	rs[1] = 0.5 * (1.0 - s) * a[1] + 0.5 * (1.0 + s) * b[1];
	drs[0] = 0.5 * (b[0] - a[0]);
	drs[1] = 0.5 * (b[1] - a[1]);
}

}  // namespace

namespace dgx {

bool build_tri_ops(int p, TriOps<double>& o)
{
	std::memset(&o, 0, sizeof o);
	TriBasis B;
	if (!B.init(p)) return false;
	const TriRule* cub = find_rule("WV", 2 * p);
	if (!cub) return false;
	std::vector<double> rs, w;
	const int Nc = expand(*cub, rs, w);
	if (Nc > kTriMaxNc || B.Np > kTriMaxNp) return false;
	const int Np = B.Np, N = p + 1;
	o.p = p;
	o.Np = Np;
	o.Nc = Nc;
	o.N = N;
	std::vector<double> phi(Np), dr(Np), ds(Np);
	for (int q = 0; q < Nc; ++q) {
		B.eval(rs[2 * q], rs[2 * q + 1], phi.data(), dr.data(), ds.data());
		o.w[q] = w[q];
		for (int m = 0; m < Np; ++m) {
			o.I[q * Np + m] = phi[m];
			o.Dr[q * Np + m] = dr[m];
			o.Ds[q * Np + m] = ds[m];
			o.DrW[m * Nc + q] = dr[m] * w[q];
			o.DsW[m * Nc + q] = ds[m] * w[q];
		}
	}
	double xg[kHpMaxN], wg[kHpMaxN];
	gauss_legendre(N, xg, wg);
	for (int lf = 0; lf < 3; ++lf)
		for (int a = 0; a < N; ++a) {
			double pt[2], dpt[2];
			face_point(lf, xg[a], pt, dpt);
			B.eval(pt[0], pt[1], phi.data(), nullptr, nullptr);
			for (int m = 0; m < Np; ++m) {
				o.Ifc[lf][a * Np + m] = phi[m];
				o.tw0[lf][m * N + a] = phi[m] * wg[a];
			}
		}
	std::vector<double> one(Nc, 1.0), mi((size_t)Np * Np), K((size_t)3 * N * N);
	if (!tri_mass_data(o, one.data(), mi.data(), K.data())) return false;
	std::copy(mi.begin(), mi.end(), o.Minv);
	for (int lf = 0; lf < 3; ++lf) std::copy(K.begin() + lf * N * N, K.begin() + (lf + 1) * N * N, o.K[lf]);
	for (int q = 0; q < Nc; ++q)
		for (int lf = 0; lf < 3; ++lf)
			for (int b = 0; b < N; ++b) {
				long double acc = 0.0L;
				for (int m = 0; m < Np; ++m) {
					long double mt = 0.0L;   // (M^-1 tw0_f)[m][b]
					for (int n = 0; n < Np; ++n) mt += (long double)mi[(size_t)m * Np + n] * o.tw0[lf][n * N + b];
					acc += (long double)o.I[q * Np + m] * mt;
				}
				o.L[q * 3 * N + lf * N + b] = (double)acc;
			}
	for (int m = 0; m < Np; ++m)
		for (int q = 0; q < Nc; ++q) {
			long double acc = 0.0L;
			for (int n = 0; n < Np; ++n) acc += (long double)mi[(size_t)m * Np + n] * o.I[q * Np + n];
			o.MIt[m * Nc + q] = (double)acc;
		}
	return true;
}

// J-scaled cofactors (y_s, -x_s, -y_r, x_r) and det J at the WV cubature of a TRI element of order p with
// geometry nodes xyz (2 x Np, the solution nodes)
void tri_geometry(const TriOps<double>& o, const double* xyz, double* met, double* detJ)
{
	const int Np = o.Np, Nc = o.Nc;
	for (int q = 0; q < Nc; ++q) {
		double xr = 0.0, xs = 0.0, yr = 0.0, ys = 0.0;
		for (int m = 0; m < Np; ++m) {
			xr += o.Dr[q * Np + m] * xyz[m];
			xs += o.Ds[q * Np + m] * xyz[m];
			yr += o.Dr[q * Np + m] * xyz[Np + m];
			ys += o.Ds[q * Np + m] * xyz[Np + m];
		}
		met[0 * Nc + q] = ys;
		met[1 * Nc + q] = -xs;
		met[2 * Nc + q] = -yr;
		met[3 * Nc + q] = xr;
		detJ[q] = xr * ys - xs * yr;
	}
}

// The element's mass matrix M = I^T diag(w det J) I (constructor_mass_T, volume_solver_T.c:115-145), its inverse
// (Np x Np, row-major) and, per local face, the BR2 lift trace K_f = Ifc_f M^-1 tw0_f (N x N): the trace at the
// element's own GL face points of the lift M^-1 tw0_f y of face values y.  false if M is singular.
bool tri_mass_data(const TriOps<double>& o, const double* detJ, double* minv, double* K)
{
	const int Np = o.Np, Nc = o.Nc, N = o.N;
	std::vector<double> M((size_t)Np * 2 * Np, 0.0);
	for (int i = 0; i < Np; ++i)
		for (int j = 0; j < Np; ++j) {
			double acc = 0.0;
			for (int q = 0; q < Nc; ++q) acc += o.I[q * Np + i] * o.w[q] * detJ[q] * o.I[q * Np + j];
			M[(size_t)i * 2 * Np + j] = acc;
		}
	for (int i = 0; i < Np; ++i) M[(size_t)i * 2 * Np + Np + i] = 1.0;
	for (int c = 0; c < Np; ++c) {   // Gauss-Jordan, partial pivoting (M is symmetric positive definite)
		int piv = c;
		for (int i = c + 1; i < Np; ++i)
			if (std::fabs(M[(size_t)i * 2 * Np + c]) > std::fabs(M[(size_t)piv * 2 * Np + c])) piv = i;
		if (!(M[(size_t)piv * 2 * Np + c] != 0.0)) return false;
		if (piv != c)
			for (int j = 0; j < 2 * Np; ++j) std::swap(M[(size_t)c * 2 * Np + j], M[(size_t)piv * 2 * Np + j]);
		const double d = 1.0 / M[(size_t)c * 2 * Np + c];
		for (int j = 0; j < 2 * Np; ++j) M[(size_t)c * 2 * Np + j] *= d;
		for (int i = 0; i < Np; ++i) {
			if (i == c) continue;
			const double f = M[(size_t)i * 2 * Np + c];
			if (f == 0.0) continue;
			for (int j = 0; j < 2 * Np; ++j) M[(size_t)i * 2 * Np + j] -= f * M[(size_t)c * 2 * Np + j];
		}
	}
	for (int i = 0; i < Np; ++i)
		for (int j = 0; j < Np; ++j) minv[i * Np + j] = M[(size_t)i * 2 * Np + Np + j];
	std::vector<double> L((size_t)Np * N);
	for (int lf = 0; lf < 3; ++lf) {
		for (int m = 0; m < Np; ++m)
			for (int b = 0; b < N; ++b) {
				double acc = 0.0;
				for (int n = 0; n < Np; ++n) acc += minv[m * Np + n] * o.tw0[lf][n * N + b];
				L[(size_t)m * N + b] = acc;
			}
		for (int a = 0; a < N; ++a)
			for (int b = 0; b < N; ++b) {
				double acc = 0.0;
				for (int m = 0; m < Np; ++m) acc += o.Ifc[lf][a * Np + m] * L[(size_t)m * N + b];
				K[(lf * N + a) * N + b] = acc;
			}
	}
	return true;
}

// outward n J_f of local face lf of a TRI element of order p at face coordinates s[0..nq) (the tangent of its map
// turned clockwise, reversed on face 1, which runs clockwise on a counter-clockwise element)
void tri_face_nj(int p, int lf, const double* xyz, const double* s, int nq, double* nj)
{
	TriBasis B;
	B.init(p);
	std::vector<double> dr(B.Np), ds(B.Np);
	for (int q = 0; q < nq; ++q) {
		double pt[2], dpt[2];
		face_point(lf, s[q], pt, dpt);
		B.eval(pt[0], pt[1], nullptr, dr.data(), ds.data());
		double tx = 0.0, ty = 0.0;
		for (int m = 0; m < B.Np; ++m) {
			const double dm = dr[m] * dpt[0] + ds[m] * dpt[1];
			tx += dm * xyz[m];
			ty += dm * xyz[B.Np + m];
		}
		const double sg = (lf == 1) ? -1.0 : 1.0;
		nj[2 * q] = sg * ty;
		nj[2 * q + 1] = -sg * tx;
	}
}

}  // namespace dgx

// Host only: the solution nodes of the reference element of `type` (0 QUAD, 1 TRI) and order p.
extern "C" dgx_status dgx_reference_nodes(int type, int p, double* rst, double* w, int* np)
{
	if (p < 1 || p > 5 || (type != 0 && type != 1)) return DGX_ERR_ARG;
	if (type == 1) {
		const TriRule* t = find_rule("WSH", p);
		if (!t) return DGX_ERR_UNSUPPORTED;
		std::vector<double> rs, ww;
		const int n = expand(*t, rs, ww);
		if (np) *np = n;
		if (rst) std::memcpy(rst, rs.data(), sizeof(double) * 2 * n);
		if (w) std::memcpy(w, ww.data(), sizeof(double) * n);
		return DGX_OK;
	}
	const int N = p + 1;
	double x[dgx::kMaxN], wl[dgx::kMaxN];
	dgx::gauss_lobatto(N, x, wl);
	if (np) *np = N * N;
	for (int j = 0; j < N; ++j)
		for (int i = 0; i < N; ++i) {
			if (rst) {
				rst[2 * (i + N * j)] = x[i];
				rst[2 * (i + N * j) + 1] = x[j];
			}
			if (w) w[i + N * j] = wl[i] * wl[j];
		}
	return DGX_OK;
}

// Host only: the cubature of the reference's TRI rule <kind><q> (kind 0: WSH, 1: WV) as the reference expands it.
extern "C" dgx_status dgx_tri_rule(int kind, int q, double* rst, double* w, int* n)
{
	if (kind != 0 && kind != 1) return DGX_ERR_ARG;
	const TriRule* t = find_rule(kind == 0 ? "WSH" : "WV", q);
	if (!t) return DGX_ERR_ARG;
	std::vector<double> rs, ww;
	const int k = expand(*t, rs, ww);
	if (k < 0) return DGX_ERR_ARG;
	if (n) *n = k;
	if (rst) std::memcpy(rst, rs.data(), sizeof(double) * 2 * k);
	if (w) std::memcpy(w, ww.data(), sizeof(double) * k);
	return DGX_OK;
}
