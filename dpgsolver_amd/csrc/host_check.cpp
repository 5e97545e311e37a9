// host_check.cpp -- host-only driver of the library's file parsers and host tables, built by `make host_check`
// with the address and undefined-behaviour sanitizers (no GPU code, no HIP runtime calls).
//
// For every gmsh file on the command line it reads the mesh at p = 1 and 3 (dgx_mesh_read_gmsh), writes the
// reference's restart file for a smooth state and reads it back (dgx_restart_write_gmsh /
// dgx_restart_read_solution), and reads truncated copies of the mesh file (which must fail with a status code,
// not crash).  It also expands every TRI node and cubature table and every face permutation, and calls each
// host entry point with arguments it must refuse.  Exit status 0 when every call behaves; the sanitizers abort
// the process on any out-of-bounds access, use after free, leak or undefined behaviour.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/dgx.h"

namespace {

int failures = 0;

void expect(bool ok, const char* what, const std::string& where)
{
	if (!ok) {
		std::fprintf(stderr, "host_check: %s failed (%s)\n", what, where.c_str());
		++failures;
	}
}

void check_tables()
{
	std::vector<double> rst(2 * 64), w(64);
	for (int p = 1; p <= 5; ++p) {
		int np = 0;
		expect(dgx_reference_nodes(0, p, rst.data(), w.data(), &np) == DGX_OK && np == (p + 1) * (p + 1),
		       "QUAD reference nodes", std::to_string(p));
		expect(dgx_reference_nodes(1, p, rst.data(), w.data(), &np) == DGX_OK && np == (p + 1) * (p + 2) / 2,
		       "TRI reference nodes", std::to_string(p));
		int n = 0;
		expect(dgx_tri_rule(0, p, rst.data(), w.data(), &n) == DGX_OK, "WSH rule", std::to_string(p));
		expect(dgx_tri_rule(1, 2 * p, rst.data(), w.data(), &n) == DGX_OK, "WV rule", std::to_string(2 * p));
	}
	expect(dgx_reference_nodes(1, 9, rst.data(), w.data(), nullptr) != DGX_OK, "TRI p = 9 refused", "");
	expect(dgx_tri_rule(2, 2, rst.data(), w.data(), nullptr) != DGX_OK, "unknown rule kind refused", "");
	std::vector<int> perm(64);
	for (int d = 2; d <= 3; ++d)
		for (int p = 1; p <= 4; ++p)
			for (int ord = -1; ord <= 9; ++ord) {
				const dgx_status st = dgx_face_permutation(d, p, ord, perm.data());
				const bool valid = d == 2 ? (ord == 0 || ord == 1) : (ord >= 0 && ord <= 7);
				expect((st == DGX_OK) == valid, "face permutation status",
				       std::to_string(d) + "d p" + std::to_string(p) + " ord " + std::to_string(ord));
			}
}

bool copy_prefix(const std::string& src, const std::string& dst, double frac)
{
	FILE* in = std::fopen(src.c_str(), "rb");
	if (!in) return false;
	std::fseek(in, 0, SEEK_END);
	const long n = std::ftell(in);
	std::fseek(in, 0, SEEK_SET);
	std::vector<char> buf((size_t)std::max(n, 1L));
	const size_t got = std::fread(buf.data(), 1, (size_t)n, in);
	std::fclose(in);
	// cut inside the part the reader needs: before the end of the $Elements section
	const std::string text(buf.data(), got);
	const size_t end = text.find("$EndElements");
	if (end == std::string::npos) return false;
	FILE* out = std::fopen(dst.c_str(), "wb");
	if (!out) return false;
	std::fwrite(buf.data(), 1, (size_t)(frac * (double)end), out);
	std::fclose(out);
	return true;
}

void check_mesh(const std::string& path, const std::string& tmpdir)
{
	for (int p : { 1, 3 }) {
		dgx_host_mesh m{};
		const dgx_status st = dgx_mesh_read_gmsh(path.c_str(), p, &m);
		if (st == DGX_ERR_UNSUPPORTED) {   // e.g. a mesh with TRI elements
			dgx_mesh_free(&m);
			continue;
		}
		expect(st == DGX_OK, "dgx_mesh_read_gmsh", path);
		if (st != DGX_OK || m.d != 2) {   // HEX meshes: read (and so checked) only; the round trip below is 2-D
			dgx_mesh_free(&m);
			continue;
		}
		const int N = p + 1, Np = N * N, nv = 4;
		std::vector<double> u((size_t)m.E * nv * Np), back(u.size());
		for (int e = 0; e < m.E; ++e)
			for (int i = 0; i < Np; ++i) {
				const double x = m.xyz[(size_t)e * 2 * Np + i], y = m.xyz[(size_t)e * 2 * Np + Np + i];
				const double rho = 1.0 + 0.1 * std::sin(x) * std::cos(y), pr = 1.0 + 0.05 * std::cos(x + y);
				const double vx = 0.3, vy = -0.1;
				double* s = &u[(size_t)e * nv * Np + i];
				s[0] = rho;
				s[Np] = rho * vx;
				s[2 * Np] = rho * vy;
				s[3 * Np] = pr / 0.4 + 0.5 * rho * (vx * vx + vy * vy);
			}
		dgx_config cfg{};
		cfg.d = 2;
		cfg.p = p;
		cfg.pde = DGX_PDE_EULER;
		cfg.num_flux = DGX_FLUX_LF;
		dgx_mesh mesh{ m.E, m.F, m.xyz, m.faces };
		const std::string rpath = tmpdir + "/host_check_restart.msh";
		expect(dgx_restart_write_gmsh(rpath.c_str(), &cfg, &mesh, u.data()) == DGX_OK, "restart write", path);
		expect(dgx_restart_read_solution(rpath.c_str(), &cfg, m.E, back.data()) == DGX_OK, "restart read", path);
		double worst = 0.0;
		for (size_t i = 0; i < u.size(); ++i) worst = std::max(worst, std::fabs(u[i] - back[i]));
		expect(worst < 1e-10, "restart round trip", path);
		expect(dgx_restart_read_solution(rpath.c_str(), &cfg, m.E + 1, back.data()) != DGX_OK,
		       "restart with the wrong element count refused", path);
		// a restart file cut inside its coefficients: a status code
		{
			FILE* in = std::fopen(rpath.c_str(), "rb");
			std::vector<char> buf;
			if (in) {
				char tmp[4096];
				size_t k;
				while ((k = std::fread(tmp, 1, sizeof tmp, in)) > 0) buf.insert(buf.end(), tmp, tmp + k);
				std::fclose(in);
			}
			const std::string text(buf.begin(), buf.end());
			const size_t at = text.find("$SolutionCoefficients");
			if (at != std::string::npos) {
				const std::string cut = tmpdir + "/host_check_restart_cut.msh";
				FILE* out = std::fopen(cut.c_str(), "wb");
				if (out) {
					std::fwrite(buf.data(), 1, at + (text.size() - at) / 2, out);
					std::fclose(out);
					expect(dgx_restart_read_solution(cut.c_str(), &cfg, m.E, back.data()) != DGX_OK,
					       "truncated restart refused", path);
					std::remove(cut.c_str());
				}
			}
		}
		std::remove(rpath.c_str());
		dgx_mesh_free(&m);
	}
	// copies truncated before the end of the elements: a status code, never a crash
	for (double frac : { 0.1, 0.37, 0.5, 0.73, 0.95, 0.98 }) {
		const std::string cut = tmpdir + "/host_check_cut.msh";
		if (!copy_prefix(path, cut, frac)) continue;
		dgx_host_mesh m{};
		expect(dgx_mesh_read_gmsh(cut.c_str(), 2, &m) != DGX_OK, "truncated mesh refused", path + " at " + std::to_string(frac));
		dgx_mesh_free(&m);
		std::remove(cut.c_str());
	}
}

}  // namespace

int main(int argc, char** argv)
{
	if (argc < 3) {
		std::fprintf(stderr, "usage: host_check TMPDIR MESH.msh...\n");
		return 2;
	}
	check_tables();
	dgx_host_mesh none{};
	expect(dgx_mesh_read_gmsh("/nonexistent/mesh.msh", 2, &none) != DGX_OK, "missing file refused", "");
	dgx_mesh_free(&none);
	for (int i = 2; i < argc; ++i) check_mesh(argv[i], argv[1]);
	if (failures) return 1;
	std::printf("host_check: ok\n");
	return 0;
}
