// dgx_restart.cpp -- the reference's restart file (host only, no HIP calls): a gmsh 2.2 mesh section followed
// by $SolutionCoefficients, one line per volume "v p n_dof n_var c...", the coefficients of the solution in
// the tensor-product Bernstein (Bezier) basis, dof fastest
// (output_restart_gmsh / fprintf_solution_bezier_gmsh, test_case/solution/restart/restart_writers.c:157-171,
//  301-412; read back by initialize_volumes_sol_coef, restart_T.c:317-420).
// The mesh section lists the element corner vertices (merged by coordinates), the boundary faces as gmsh line
// or quadrangle elements tagged with the reference's boundary codes (10000 + 101 Riemann, 10000 + 102 slip
// wall), then the volumes, with two tags each, as fprintf_elements_gmsh writes them.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "dgx_internal.h"

namespace {

// nodal (GLL Lagrange) <-> Bezier along every direction: M is B^-1 (to Bezier) or B (to nodal), N x N
void tensor_apply(int d, int N, const double* M, const double* in, double* out)
{
	const int Np = (d == 2) ? N * N : N * N * N;
	std::vector<double> a(in, in + Np), b(Np);
	for (int dir = 0; dir < d; ++dir) {
		const int S = (dir == 0) ? 1 : (dir == 1 ? N : N * N);
		for (int n = 0; n < Np; ++n) {
			const int c = (n / S) % N, base = n - c * S;
			double acc = 0.0;
			for (int j = 0; j < N; ++j) acc += M[c * N + j] * a[base + j * S];
			b[n] = acc;
		}
		a.swap(b);
	}
	std::memcpy(out, a.data(), sizeof(double) * Np);
}

int nv_of(const dgx_config* cfg) { return cfg->pde == DGX_PDE_ADVECTION ? 1 : cfg->d + 2; }

}  // namespace

extern "C" {

dgx_status dgx_restart_write_gmsh(const char* path, const dgx_config* cfg, const dgx_mesh* mesh, const double* u)
{
	if (!path || !cfg || !mesh || !u || !mesh->xyz || !mesh->faces || mesh->E <= 0) return DGX_ERR_ARG;
	const int d = cfg->d, p = cfg->p, N = p + 1, nv = nv_of(cfg);
	if ((d != 2 && d != 3) || p < 1 || p + 1 > dgx::kMaxN) return DGX_ERR_ARG;
	const int Np = (d == 2) ? N * N : N * N * N;
	dgx::Ops1D<double> o;
	dgx::Ops1D<float> o32;
	dgx::build_ops(d, p, o, o32);

	// corner vertices in tensor-product order, merged by coordinates
	const int nc = 1 << d;
	std::vector<int> corner(nc);
	for (int k = 0; k < nc; ++k) {
		const int i = (k & 1) ? N - 1 : 0, j = (k & 2) ? N - 1 : 0, l = (k & 4) ? N - 1 : 0;
		corner[k] = i + N * j + N * N * l;
	}
	std::map<std::vector<long long>, int> ids;
	std::vector<std::vector<double>> verts;
	std::vector<int> vnum((size_t)mesh->E * nc);
	for (int e = 0; e < mesh->E; ++e)
		for (int k = 0; k < nc; ++k) {
			std::vector<double> x(d);
			std::vector<long long> key(d);
			for (int dd = 0; dd < d; ++dd) {
				x[dd] = mesh->xyz[((size_t)e * d + dd) * Np + corner[k]];
				key[dd] = std::llround(x[dd] * 1e10);
			}
			auto it = ids.find(key);
			if (it == ids.end()) {
				it = ids.emplace(key, (int)verts.size() + 1).first;
				verts.push_back(x);
			}
			vnum[(size_t)e * nc + k] = it->second;
		}

	FILE* f = std::fopen(path, "w");
	if (!f) return DGX_ERR_ARG;
	std::fprintf(f, "$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n%zu\n", verts.size());
	for (size_t n = 0; n < verts.size(); ++n) {
		std::fprintf(f, "%zu", n + 1);
		for (int dd = 0; dd < 3; ++dd) std::fprintf(f, " %.17g", dd < d ? verts[n][dd] : 0.0);
		std::fprintf(f, "\n");
	}
	std::fprintf(f, "$EndNodes\n");
	// boundary faces (2-D: lines, 3-D: quadrangles) then volumes (quadrangles / hexahedra)
	std::vector<std::string> rows;
	for (int fc = 0; fc < mesh->F; ++fc) {
		const int* r = mesh->faces + 6 * fc;
		if (r[2] >= 0) continue;
		const int e = r[0], lf = r[1], dir = lf >> 1, side = lf & 1;
		if (e < 0 || e >= mesh->E || lf < 0 || lf >= 2 * d) {   // a face table of another dimension or size
			std::fclose(f);
			return DGX_ERR_ARG;
		}
		const int tag = 10000 + (r[5] == DGX_BC_SLIPWALL ? 102 : 101);
		std::vector<int> fvs;
		for (int k = 0; k < nc; ++k)
			if (((k >> dir) & 1) == side) fvs.push_back(vnum[(size_t)e * nc + k]);
		char buf[160];
		if (d == 2)
			std::snprintf(buf, sizeof buf, "1 2 %d 0 %d %d", tag, fvs[0], fvs[1]);
		else   // tensor-product order (a, b, d, c) -> gmsh counter-clockwise
			std::snprintf(buf, sizeof buf, "3 2 %d 0 %d %d %d %d", tag, fvs[0], fvs[1], fvs[3], fvs[2]);
		rows.emplace_back(buf);
	}
	for (int e = 0; e < mesh->E; ++e) {
		const int* v = &vnum[(size_t)e * nc];
		char buf[200];
		if (d == 2)
			std::snprintf(buf, sizeof buf, "3 2 0 0 %d %d %d %d", v[0], v[1], v[3], v[2]);
		else
			std::snprintf(buf, sizeof buf, "5 2 0 0 %d %d %d %d %d %d %d %d", v[0], v[1], v[3], v[2], v[4], v[5], v[7],
			              v[6]);
		rows.emplace_back(buf);
	}
	std::fprintf(f, "$Elements\n%zu\n", rows.size());
	for (size_t i = 0; i < rows.size(); ++i) std::fprintf(f, "%zu %s\n", i + 1, rows[i].c_str());
	std::fprintf(f, "$EndElements\n$SolutionCoefficients\n%d\n", mesh->E);
	std::vector<double> c(Np);
	for (int e = 0; e < mesh->E; ++e) {
		std::fprintf(f, "%d %d %d %d", e + 1, p, Np, nv);
		for (int v = 0; v < nv; ++v) {
			tensor_apply(d, N, o.Binv, u + ((size_t)e * nv + v) * Np, c.data());
			for (int i = 0; i < Np; ++i) std::fprintf(f, " % .15e", c[i]);
		}
		std::fprintf(f, "\n");
	}
	std::fprintf(f, "$EndSolutionCoefficients\n");
	return std::fclose(f) == 0 ? DGX_OK : DGX_ERR_ARG;
}

dgx_status dgx_restart_read_solution(const char* path, const dgx_config* cfg, int E, double* u)
{
	if (!path || !cfg || !u || E <= 0) return DGX_ERR_ARG;
	const int d = cfg->d, p = cfg->p, N = p + 1, nv = nv_of(cfg);
	if ((d != 2 && d != 3) || p < 1 || p + 1 > dgx::kMaxN) return DGX_ERR_ARG;
	const int Np = (d == 2) ? N * N : N * N * N;
	dgx::Ops1D<double> o;
	dgx::Ops1D<float> o32;
	dgx::build_ops(d, p, o, o32);
	FILE* f = std::fopen(path, "r");
	if (!f) return DGX_ERR_ARG;
	char line[256];
	bool found = false;
	while (std::fgets(line, sizeof line, f))
		if (std::strstr(line, "$SolutionCoefficients")) { found = true; break; }
	int nvol = 0;
	dgx_status st = DGX_OK;
	if (!found || std::fscanf(f, "%d", &nvol) != 1 || nvol != E) st = DGX_ERR_ARG;
	std::vector<double> c(Np);
	for (int e = 0; e < E && st == DGX_OK; ++e) {
		int id, pp, ndof, nvar;
		if (std::fscanf(f, "%d %d %d %d", &id, &pp, &ndof, &nvar) != 4 || id != e + 1 || pp != p || ndof != Np ||
		    nvar != nv) {
			st = (pp != p && ndof > 0) ? DGX_ERR_UNSUPPORTED : DGX_ERR_ARG;
			break;
		}
		for (int v = 0; v < nv && st == DGX_OK; ++v) {
			for (int i = 0; i < Np; ++i)
				if (std::fscanf(f, "%lf", &c[i]) != 1) { st = DGX_ERR_ARG; break; }
			if (st == DGX_OK) tensor_apply(d, N, o.B, c.data(), u + ((size_t)e * nv + v) * Np);
		}
	}
	std::fclose(f);
	return st;
}

}  // extern "C"
