// dgx_ops.cpp -- host construction of the 1-D operator tables (FP64 master, FP32 copy).
//
// Nodes: Gauss-Legendre cubature and Gauss-Lobatto solution nodes, p+1 per direction
// (cubature choice src/element/nodes_operators.c:314-356, 589-610; GLL interpolation
// nodes per the interp_tp GLL control setting).  Lagrange basis on GLL nodes
// (basis_sol lagrange), Bernstein basis for the positivity limiter (solve.c:187-233).
#include <cmath>
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "dgx_internal.h"

namespace dgx {

static void legendre(int n, double x, double& P, double& dP)
{
	double p0 = 1.0, p1 = x;
	if (n == 0) { P = 1.0; dP = 0.0; return; }
	for (int k = 2; k <= n; ++k) {
		const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
		p0 = p1;
		p1 = p2;
	}
	P = p1;
	dP = n * (x * p1 - p0) / (x * x - 1.0);
}

void gauss_legendre(int n, double* x, double* w)
{
	for (int i = 0; i < n; ++i) {
		double z = -std::cos(M_PI * (i + 0.75) / (n + 0.5)), P = 0, dP = 1;
		for (int it = 0; it < 100; ++it) {
			legendre(n, z, P, dP);
			const double dz = P / dP;
			z -= dz;
			if (std::fabs(dz) < 1e-16) break;
		}
		legendre(n, z, P, dP);
		x[i] = z;
		w[i] = 2.0 / ((1.0 - z * z) * dP * dP);
	}
	for (int i = 0; i < n / 2; ++i) {
		const double a = 0.5 * (x[n - 1 - i] - x[i]), b = 0.5 * (w[i] + w[n - 1 - i]);
		x[i] = -a; x[n - 1 - i] = a; w[i] = b; w[n - 1 - i] = b;
	}
	if (n % 2) x[n / 2] = 0.0;
}

void gauss_lobatto(int n, double* x, double* w)
{
	const int N = n - 1;
	x[0] = -1.0;
	x[n - 1] = 1.0;
	for (int i = 1; i < N; ++i) {
		double z = -std::cos(M_PI * i / N);
		for (int it = 0; it < 100; ++it) {
			double P, dP, Pm, dPm, Pp, dPp;
			legendre(N - 1, z, Pm, dPm);
			legendre(N + 1, z, Pp, dPp);
			(void)P; (void)dP;
			const double dz = (Pp - Pm) / (dPp - dPm);
			z -= dz;
			if (std::fabs(dz) < 1e-16) break;
		}
		x[i] = z;
	}
	for (int i = 0; i < n; ++i) {
		double P, dP;
		if (std::fabs(x[i]) < 1.0) legendre(N, x[i], P, dP);
		else P = (x[i] > 0 || N % 2 == 0) ? 1.0 : -1.0;
		w[i] = 2.0 / (N * (N + 1.0) * P * P);
	}
	for (int i = 0; i < n / 2; ++i) {
		const double a = 0.5 * (x[n - 1 - i] - x[i]), b = 0.5 * (w[i] + w[n - 1 - i]);
		x[i] = -a; x[n - 1 - i] = a; w[i] = b; w[n - 1 - i] = b;
	}
	if (n % 2) x[n / 2] = 0.0;
}

// l_c(x) and l_c'(x) of the Lagrange basis on xin
static void lagrange(int nin, const double* xin, int nout, const double* xout, double* Iop, double* Dop)
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

static bool invert(int n, const double* A, double* Ainv)
{
	std::vector<double> M(2 * n * n);
	for (int i = 0; i < n; ++i)
		for (int j = 0; j < 2 * n; ++j) M[i * 2 * n + j] = (j < n ? A[i * n + j] : (j - n == i ? 1.0 : 0.0));
	for (int col = 0; col < n; ++col) {
		int piv = col;
		for (int i = col + 1; i < n; ++i)
			if (std::fabs(M[i * 2 * n + col]) > std::fabs(M[piv * 2 * n + col])) piv = i;
		if (M[piv * 2 * n + col] == 0.0) return false;
		if (piv != col)
			for (int j = 0; j < 2 * n; ++j) std::swap(M[col * 2 * n + j], M[piv * 2 * n + j]);
		const double inv = 1.0 / M[col * 2 * n + col];
		for (int j = 0; j < 2 * n; ++j) M[col * 2 * n + j] *= inv;
		for (int i = 0; i < n; ++i) {
			if (i == col) continue;
			const double f = M[i * 2 * n + col];
			for (int j = 0; j < 2 * n; ++j) M[i * 2 * n + j] -= f * M[col * 2 * n + j];
		}
	}
	for (int i = 0; i < n; ++i)
		for (int j = 0; j < n; ++j) Ainv[i * n + j] = M[i * 2 * n + n + j];
	return true;
}

void build_ops(int d, int p, Ops1D<double>& o, Ops1D<float>& o32)
{
	std::memset(&o, 0, sizeof(o));
	const int N = p + 1;
	double xl[kMaxN], wl[kMaxN], xg[kMaxN], wg[kMaxN];
	gauss_lobatto(N, xl, wl);
	gauss_legendre(N, xg, wg);
	double I[kMaxN * kMaxN], D1[kMaxN * kMaxN], Dh[kMaxN * kMaxN], Dg[kMaxN * kMaxN];
	lagrange(N, xl, N, xg, I, D1);
	lagrange(N, xg, N, xg, nullptr, Dh);
	lagrange(N, xl, N, xl, nullptr, Dg);
	double Iinv[kMaxN * kMaxN];
	invert(N, I, Iinv);
	for (int r = 0; r < N; ++r)
		for (int c = 0; c < N; ++c) {
			o.I[r * N + c] = I[r * N + c];
			o.IT[r * N + c] = I[c * N + r];
			o.DhT[r * N + c] = Dh[c * N + r];
			o.Iinv[r * N + c] = Iinv[r * N + c];
			o.Dg[r * N + c] = Dg[r * N + c];
		}
	// Bernstein at GLL nodes
	for (int r = 0; r < N; ++r) {
		const double a = 0.5 * (1.0 + xl[r]), b = 0.5 * (1.0 - xl[r]);
		for (int j = 0; j < N; ++j) {
			double binom = 1.0;
			for (int k = 1; k <= j; ++k) binom = binom * (p - j + k) / k;
			o.B[r * N + j] = binom * std::pow(a, j) * std::pow(b, p - j);
		}
	}
	double Bt[kMaxN * kMaxN], Binv[kMaxN * kMaxN];
	for (int i = 0; i < N * N; ++i) Bt[i] = o.B[i];
	invert(N, Bt, Binv);
	for (int i = 0; i < N * N; ++i) o.Binv[i] = Binv[i];
	for (int i = 0; i < N; ++i) {
		o.Ext[0][i] = (i == 0);
		o.Ext[1][i] = (i == N - 1);
		o.w[i] = wg[i];
		o.wgll[i] = wl[i];
		// Lift[side] = I^{-T} e_side  (rhs face lift = Ibar^T (Lift (x) g))
		o.Lift[0][i] = Iinv[0 * N + i];
		o.Lift[1][i] = Iinv[(N - 1) * N + i];
	}
	// orientations: bit0 transpose, bit1 flip first coordinate, bit2 flip second
	const int NF = (d == 3) ? N * N : N;
	for (int ord = 0; ord < 8; ++ord) {
		for (int i = 0; i < NF; ++i) {
			int a = i % N, b = (d == 3) ? i / N : 0;
			int ap = a, bp = b;
			if (d == 3 && (ord & 1)) { ap = b; bp = a; }
			if (ord & 2) ap = N - 1 - ap;
			if (d == 3 && (ord & 4)) bp = N - 1 - bp;
			o.perm[ord][i] = ap + N * bp;
		}
	}
	// FP32 copy
	auto cp = [](const double* s, float* t, int n) { for (int i = 0; i < n; ++i) t[i] = (float)s[i]; };
	const int M2 = kMaxN * kMaxN;
	cp(o.I, o32.I, M2); cp(o.IT, o32.IT, M2); cp(o.DhT, o32.DhT, M2); cp(o.Iinv, o32.Iinv, M2);
	cp(o.Dg, o32.Dg, M2); cp(o.B, o32.B, M2); cp(o.Binv, o32.Binv, M2);
	cp(o.Ext[0], o32.Ext[0], 2 * kMaxN); cp(o.Lift[0], o32.Lift[0], 2 * kMaxN);
	cp(o.w, o32.w, kMaxN); cp(o.wgll, o32.wgll, kMaxN);
	std::memcpy(o32.perm, o.perm, sizeof(o.perm));
}

// Z-order (Morton) sort of element centroids: consecutive positions are spatial neighbours, so
// an XCD walking a contiguous range of positions re-reads neighbour traces from its own L2.
std::vector<int> morton_order(int d, int E, int Np, const double* xyz)
{
	std::vector<double> c((size_t)E * d);
	double lo[3] = { 1e300, 1e300, 1e300 }, hi[3] = { -1e300, -1e300, -1e300 };
	for (int e = 0; e < E; ++e)
		for (int k = 0; k < d; ++k) {
			double s = 0.0;
			for (int n = 0; n < Np; ++n) s += xyz[((size_t)e * d + k) * Np + n];
			s /= Np;
			c[(size_t)e * d + k] = s;
			lo[k] = std::min(lo[k], s);
			hi[k] = std::max(hi[k], s);
		}
	const int bits = (d == 3) ? 21 : 31;
	const double scale = (double)((1u << (bits > 30 ? 30 : bits)) - 1);
	std::vector<std::pair<uint64_t, int>> key((size_t)E);
	for (int e = 0; e < E; ++e) {
		uint64_t q[3] = { 0, 0, 0 };
		for (int k = 0; k < d; ++k) {
			const double span = hi[k] - lo[k];
			const double t = span > 0 ? (c[(size_t)e * d + k] - lo[k]) / span : 0.0;
			q[k] = (uint64_t)(t * scale + 0.5);
		}
		uint64_t m = 0;
		const int nb = (d == 3) ? 21 : 30;
		for (int b = nb - 1; b >= 0; --b)
			for (int k = d - 1; k >= 0; --k) m = (m << 1) | ((q[k] >> b) & 1u);
		key[(size_t)e] = { m, e };
	}
	std::stable_sort(key.begin(), key.end());
	std::vector<int> order((size_t)E);
	for (int i = 0; i < E; ++i) order[(size_t)i] = key[(size_t)i].second;
	return order;
}

// minimum straight edge (vertex-to-vertex) length per element (compute_min_length_measure,
// solve_dg.c:323-379); TP vertices are the GLL corner nodes
std::vector<double> min_edge_lengths(int d, int N, int E, const double* xyz)
{
	const int Np = (d == 3) ? N * N * N : N * N;
	auto vnode = [&](int v) {
		const int b0 = v & 1, b1 = (v >> 1) & 1, b2 = (v >> 2) & 1;
		return b0 * (N - 1) + N * (b1 * (N - 1) + N * (b2 * (N - 1)));
	};
	std::vector<double> out((size_t)E);
	for (int e = 0; e < E; ++e) {
		double m = 1e300;
		const int nv = 1 << d;
		for (int a = 0; a < nv; ++a)
			for (int k = 0; k < d; ++k) {
				const int b = a ^ (1 << k);
				if (b < a) continue;  // each edge once: vertices differing in one bit
				double s = 0.0;
				for (int x = 0; x < d; ++x) {
					const double t = xyz[((size_t)e * d + x) * Np + vnode(a)] - xyz[((size_t)e * d + x) * Np + vnode(b)];
					s += t * t;
				}
				m = std::min(m, std::sqrt(s));
			}
		out[(size_t)e] = m;
	}
	return out;
}

}  // namespace dgx

// Host-only: the face node permutation the library applies for an orientation code at the ABI.
extern "C" dgx_status dgx_face_permutation(int d, int p, int ord, int* perm)
{
	if ((d != 2 && d != 3) || p < 1 || p > dgx::kMaxN - 1 || !perm) return DGX_ERR_ARG;
	const int bits = dgx::ord_bits(d, ord);
	if (bits < 0) return DGX_ERR_ARG;
	dgx::Ops1D<double> o;
	dgx::Ops1D<float> o32;
	dgx::build_ops(d, p, o, o32);
	const int N = p + 1, NF = (d == 3) ? N * N : N;
	for (int i = 0; i < NF; ++i) perm[i] = o.perm[bits][i];
	return DGX_OK;
}
