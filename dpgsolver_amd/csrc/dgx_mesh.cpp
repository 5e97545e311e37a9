// dgx_mesh.cpp -- gmsh 2.2 ASCII mesh ingest for 2-D QUAD, TRI and mixed meshes and 3-D HEX meshes (host only,
// no HIP calls).
//
// Follows the reference's reader and connectivity (src/simulation/mesh/mesh_readers.c,
// mesh_connectivity.c:285-465, mesh_periodic.c):
//  * volumes are the 2-D elements in file order; gmsh's counter-clockwise QUAD vertices (a, b, c, d) are
//    taken in tensor-product order (a, b, d, c), r fastest; TRI vertices (a, b, c) as they are;
//  * QUAD local faces 0..3 are r = -1, r = +1, s = -1, s = +1 with their vertices in increasing face coordinate;
//    TRI local face f is the edge opposite vertex f, vertices (1, 2), (0, 2), (0, 1);
//  * periodic faces are the boundary line elements with a periodic tag (x051..x056 translation, x061/x062
//    reflected in y); masters (x051, x053, x055, x061) and slaves (x052, x054, x056, x062) are paired in the
//    order of their centroids along the face (the slave's reflected for x062), and a volume face lying on a
//    slave takes the master's vertices (correct_f_ve_for_periodic, mesh_periodic.c:102-135, 240-440);
//  * faces match when their vertex sets then agree; v_to_lf_wp keeps only matches whose raw vertex sets agree
//    and holds the periodic boundary tags elsewhere;
//  * unmatched faces take the physical tag of the boundary line element with the same vertices (add_bc_info).
// The geometry is the straight-sided map of the vertices (QUAD bilinear, TRI affine) evaluated at the element's
// solution nodes of degree p (QUAD: GLL tensor nodes; TRI: the WSH nodes of dgx_reference_nodes).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/dgx.h"

namespace {

struct Elem {
	int type = 0;
	std::vector<int> tags;
	std::vector<int> nodes;   // gmsh node ids
};

void gll_nodes(int n, std::vector<double>& x)
{
	// Gauss-Lobatto-Legendre nodes on [-1, 1]: the end points and the roots of P'_{n-1}
	x.assign(n, 0.0);
	x[0] = -1.0;
	x[n - 1] = 1.0;
	const int m = n - 1;
	for (int i = 1; i < m; ++i) {
		double z = -std::cos(M_PI * i / m);
		for (int it = 0; it < 100; ++it) {
			// P_m and its first two derivatives by the three-term recurrence
			double p0 = 1.0, p1 = z;
			for (int k = 2; k <= m; ++k) {
				const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
				p0 = p1;
				p1 = p2;
			}
			const double dp = m * (z * p1 - p0) / (z * z - 1.0);
			const double d2p = (2.0 * z * dp - m * (m + 1.0) * p1) / (1.0 - z * z);
			const double dz = dp / d2p;
			z -= dz;
			if (std::fabs(dz) < 1e-16) break;
		}
		x[i] = z;
	}
	std::sort(x.begin(), x.end());
}

int bc_code(int tag)
{
	switch (tag % 10000) {
	case 101: return DGX_BC_RIEMANN;
	case 102: return DGX_BC_SLIPWALL;
	case 111: return DGX_BC_NOSLIP_ADIABATIC;
	case 112: return DGX_BC_NOSLIP_DIABATIC;
	case 113: return DGX_BC_NOSLIP_ROTATING;
	default: return -1;
	}
}

}  // namespace


// what the readers return, before it is copied into the C structures
struct ReadOut {
	std::vector<int> type;                      // per volume: 0 QUAD, 1 TRI
	std::vector<double> xyz;                    // per volume 2 x Np, concatenated
	std::vector<int> faces, v2v, v2lf, v2lfwp;  // v2*: 4 per volume (a TRI's 4th is -1)
	int n_unsupported = 0, n_inverted = 0;
};

static dgx_status read_any(const char* path, int p, bool allow_tri, ReadOut& R)
{
	std::ifstream in(path);
	if (!in) return DGX_ERR_ARG;
	std::map<int, std::array<double, 2>> nodes;
	std::vector<Elem> elems;
	std::string line;
	while (std::getline(in, line)) {
		if (line.rfind("$MeshFormat", 0) == 0) {
			std::getline(in, line);
			std::istringstream ss(line);
			double ver = 0;
			int ft = 1;
			ss >> ver >> ft;
			if (ver < 2.0 || ver >= 3.0 || ft != 0) return DGX_ERR_UNSUPPORTED;   // gmsh 2.x ASCII only
		} else if (line.rfind("$Nodes", 0) == 0) {
			int n = 0;
			in >> n;
			for (int i = 0; i < n; ++i) {
				int id = 0;
				double x = 0.0, y = 0.0, z = 0.0;
				if (!(in >> id >> x >> y >> z)) return DGX_ERR_ARG;   // truncated or malformed node line
				nodes[id] = { x, y };
			}
		} else if (line.rfind("$Elements", 0) == 0) {
			int n = 0;
			in >> n;
			std::getline(in, line);
			for (int i = 0; i < n; ++i) {
				std::getline(in, line);
				std::istringstream ss(line);
				int id = 0, type = 0, ntag = -1;
				if (!(ss >> id >> type >> ntag) || ntag < 0 || ntag > 64) return DGX_ERR_ARG;
				Elem e;
				e.type = type;
				e.tags.resize(ntag);
				for (int t = 0; t < ntag; ++t)
					if (!(ss >> e.tags[t])) return DGX_ERR_ARG;
				int v;
				while (ss >> v) e.nodes.push_back(v);
				// vertices per gmsh element type: 15 point, 1 line, 2 triangle, 3 quadrangle, 4 tetrahedron,
				// 5 hexahedron (first-order types only)
				const int need = type == 15 ? 1 : type == 1 ? 2 : type == 2 ? 3 : type == 3 ? 4 : type == 4 ? 4 : type == 5 ? 8 : -1;
				if (need < 0) return DGX_ERR_UNSUPPORTED;
				if ((int)e.nodes.size() < need) return DGX_ERR_ARG;   // truncated element line
				for (int k = 0; k < need; ++k)
					if (!nodes.count(e.nodes[k])) return DGX_ERR_ARG;   // a vertex the $Nodes section lacks
				elems.push_back(e);
			}
		}
	}
	if (nodes.empty() || elems.empty()) return DGX_ERR_ARG;
	for (const Elem& e : elems)
		if ((e.type == 2 && !allow_tri) || e.type == 4 || e.type == 5) return DGX_ERR_UNSUPPORTED;   // TRI / 3-D
	auto key = [](std::array<int, 2> x) { return std::make_pair(std::min(x[0], x[1]), std::max(x[0], x[1])); };

	// periodic boundary faces: pair masters and slaves by (direction, centroid along the face)
	struct PFace {
		int dir;
		double c;
		std::array<int, 2> v;
		int tag;
	};
	std::vector<PFace> pm, ps;
	for (const Elem& e : elems) {
		if (e.type != 1 || e.tags.empty()) continue;
		const int base = e.tags[0] % 10000;
		const bool m = (base == 51 || base == 53 || base == 55 || base == 61);
		const bool sl = (base == 52 || base == 54 || base == 56 || base == 62);
		if (!m && !sl) continue;
		if (base == 55 || base == 56) return DGX_ERR_UNSUPPORTED;   // z-periodic: 3-D only
		PFace f;
		f.dir = (base == 53 || base == 54) ? 1 : 0;
		const int t = 1 - f.dir;   // the coordinate along the face
		const double scale = (sl && base == 62) ? -1.0 : 1.0;
		f.c = 0.5 * scale * (nodes[e.nodes[0]][t] + nodes[e.nodes[1]][t]);
		f.v = { e.nodes[0], e.nodes[1] };
		f.tag = e.tags[0];
		(m ? pm : ps).push_back(f);
	}
	if (pm.size() != ps.size()) return DGX_ERR_ARG;
	auto by_pos = [](const PFace& a, const PFace& b) { return a.dir != b.dir ? a.dir < b.dir : a.c < b.c; };
	std::sort(pm.begin(), pm.end(), by_pos);
	std::sort(ps.begin(), ps.end(), by_pos);
	// slave face (raw vertex pair) -> its vertices' master counterparts, matched along the face
	std::map<std::pair<int, int>, std::map<int, int>> slave_map;
	for (size_t i = 0; i < ps.size(); ++i) {
		const PFace &M = pm[i], &S = ps[i];
		if (M.dir != S.dir) return DGX_ERR_ARG;
		const int t = 1 - S.dir;
		const double sc = (S.tag % 10000 == 62) ? -1.0 : 1.0;
		const double s0 = sc * nodes[S.v[0]][t], m0 = nodes[M.v[0]][t], m1 = nodes[M.v[1]][t];
		std::map<int, int> mp;
		if (std::fabs(s0 - m0) <= std::fabs(s0 - m1)) { mp[S.v[0]] = M.v[0]; mp[S.v[1]] = M.v[1]; }
		else { mp[S.v[0]] = M.v[1]; mp[S.v[1]] = M.v[0]; }
		slave_map[key(S.v)] = mp;
	}

	// volumes and their local faces (four face slots per volume; a TRI uses three)
	std::vector<std::array<int, 4>> vol;   // QUAD: tensor-product vertex order; TRI: (a, b, c, -1)
	std::vector<int> vtype;
	for (const Elem& e : elems) {
		if (e.type == 3) {
			vol.push_back({ e.nodes[0], e.nodes[1], e.nodes[3], e.nodes[2] });
			vtype.push_back(0);
		} else if (e.type == 2) {
			vol.push_back({ e.nodes[0], e.nodes[1], e.nodes[2], -1 });
			vtype.push_back(1);
		}
	}
	const int E = (int)vol.size();
	if (E == 0) return DGX_ERR_UNSUPPORTED;
	static const int fvq[4][2] = { { 0, 2 }, { 1, 3 }, { 0, 1 }, { 2, 3 } };
	static const int fvt[3][2] = { { 1, 2 }, { 0, 2 }, { 0, 1 } };
	auto nfaces = [&](int v) { return vtype[v] ? 3 : 4; };
	const int nf = 4 * E;
	std::vector<std::array<int, 2>> raw(nf, { -1, -1 }), cor(nf, { -1, -1 });
	for (int v = 0; v < E; ++v)
		for (int lf = 0; lf < nfaces(v); ++lf) {
			const int* fv = vtype[v] ? fvt[lf] : fvq[lf];
			const int a = vol[v][fv[0]], b = vol[v][fv[1]];
			raw[4 * v + lf] = { a, b };
			auto it = slave_map.find(key({ a, b }));
			if (it == slave_map.end()) cor[4 * v + lf] = { a, b };
			else cor[4 * v + lf] = { it->second.at(a), it->second.at(b) };
		}
	std::map<std::pair<int, int>, std::vector<int>> groups;
	for (int f = 0; f < nf; ++f)
		if (f % 4 < nfaces(f / 4)) groups[key(cor[f])].push_back(f);
	std::map<std::pair<int, int>, int> btag;   // boundary line elements by raw vertex pair
	for (const Elem& e : elems)
		if (e.type == 1 && !e.tags.empty()) btag[key({ e.nodes[0], e.nodes[1] })] = e.tags[0];

	std::vector<int> v2v(nf, -1), v2lf(nf, -1), v2lfwp(nf, -1);
	std::vector<int> faces;
	int n_unsupported = 0;
	for (const auto& g : groups) {
		const std::vector<int>& fs = g.second;
		if (fs.size() == 2) {
			const int f0 = std::min(fs[0], fs[1]), f1 = std::max(fs[0], fs[1]);
			v2v[f0] = f1 / 4;
			v2lf[f0] = f1 % 4;
			v2v[f1] = f0 / 4;
			v2lf[f1] = f0 % 4;
			if (key(raw[f0]) == key(raw[f1])) {
				v2lfwp[f0] = v2lf[f0];
				v2lfwp[f1] = v2lf[f1];
			}
			// same direction along the face (after the periodic map) or reversed
			const int ord = (cor[f0][0] == cor[f1][0]) ? 0 : 1;   // ind_ord: LINE face aligned or reversed
			const int row[6] = { f0 / 4, f0 % 4, f1 / 4, f1 % 4, ord, DGX_BC_INTERIOR };
			faces.insert(faces.end(), row, row + 6);
		} else if (fs.size() == 1) {
			const int f = fs[0];
			auto it = btag.find(key(raw[f]));
			const int tag = (it == btag.end()) ? -1 : it->second;
			v2lf[f] = tag;
			const int bc = bc_code(tag);
			if (bc < 0) ++n_unsupported;
			const int row[6] = { f / 4, f % 4, -1, -1, 0, bc < 0 ? DGX_BC_RIEMANN : bc };
			faces.insert(faces.end(), row, row + 6);
		} else {
			return DGX_ERR_UNSUPPORTED;   // non-manifold face
		}
	}
	// periodic boundary tags in v_to_lf_wp (add_bc_info with include_periodic)
	for (int f = 0; f < nf; ++f)
		if (f % 4 < nfaces(f / 4) && v2lfwp[f] == -1) {
			auto it = btag.find(key(raw[f]));
			if (it != btag.end()) v2lfwp[f] = it->second;
		}

	// straight-sided geometry at the solution nodes of degree p
	const int N = p + 1, NpQ = N * N, NpT = N * (N + 1) / 2;
	std::vector<double> r;
	gll_nodes(N, r);
	std::vector<double> rst(2 * (size_t)NpT);
	if (allow_tri && dgx_reference_nodes(1, p, rst.data(), nullptr, nullptr) != DGX_OK) return DGX_ERR_UNSUPPORTED;
	int n_inverted = 0;
	for (int v = 0; v < E; ++v) {
		double X[4][2] = {};
		for (int k = 0; k < (vtype[v] ? 3 : 4); ++k) {
			X[k][0] = nodes[vol[v][k]][0];
			X[k][1] = nodes[vol[v][k]][1];
		}
		const size_t base = R.xyz.size();
		if (vtype[v] == 0) {
			R.xyz.resize(base + 2 * (size_t)NpQ);
			for (int j2 = 0; j2 < N; ++j2)
				for (int i2 = 0; i2 < N; ++i2) {
					const double s0 = 0.5 * (1 - r[i2]), s1 = 0.5 * (1 + r[i2]), t0 = 0.5 * (1 - r[j2]), t1 = 0.5 * (1 + r[j2]);
					for (int dd = 0; dd < 2; ++dd)
						R.xyz[base + (size_t)dd * NpQ + i2 + N * j2] =
						    s0 * t0 * X[0][dd] + s1 * t0 * X[1][dd] + s0 * t1 * X[2][dd] + s1 * t1 * X[3][dd];
				}
			// orientation: det J at the first vertex of the bilinear map
			const double j00 = X[1][0] - X[0][0], j01 = X[2][0] - X[0][0], j10 = X[1][1] - X[0][1], j11 = X[2][1] - X[0][1];
			if (j00 * j11 - j01 * j10 <= 0.0) ++n_inverted;
		} else {
			// barycentric coordinates on the reference triangle (-1, -1/sqrt3), (1, -1/sqrt3), (0, 2/sqrt3)
			R.xyz.resize(base + 2 * (size_t)NpT);
			const double s3 = std::sqrt(3.0);
			for (int k = 0; k < NpT; ++k) {
				const double l2 = (rst[2 * k + 1] + 1.0 / s3) / s3;
				const double l1 = 0.5 * ((rst[2 * k] + 1.0) - l2);
				const double l0 = 1.0 - l1 - l2;
				for (int dd = 0; dd < 2; ++dd)
					R.xyz[base + (size_t)dd * NpT + k] = l0 * X[0][dd] + l1 * X[1][dd] + l2 * X[2][dd];
			}
			const double cr = (X[1][0] - X[0][0]) * (X[2][1] - X[0][1]) - (X[1][1] - X[0][1]) * (X[2][0] - X[0][0]);
			if (cr <= 0.0) ++n_inverted;
		}
	}
	R.type = vtype;
	R.faces = faces;
	R.v2v = v2v;
	R.v2lf = v2lf;
	R.v2lfwp = v2lfwp;
	for (int v = 0; v < E; ++v)
		if (vtype[v]) R.v2v[4 * v + 3] = R.v2lf[4 * v + 3] = R.v2lfwp[4 * v + 3] = -1;
	R.n_unsupported = n_unsupported;
	R.n_inverted = n_inverted;
	return DGX_OK;
}

template <typename T>
static T* dup(const std::vector<T>& v)
{
	T* out = static_cast<T*>(std::malloc(sizeof(T) * std::max<size_t>(v.size(), 1)));
	if (out) std::copy(v.begin(), v.end(), out);
	return out;
}

static dgx_status read_gmsh(const char* path, int p, dgx_host_mesh* out)
{
	ReadOut R;
	const dgx_status st = read_any(path, p, false, R);
	if (st != DGX_OK) return st;
	const int E = (int)R.type.size();
	out->d = 2;
	out->p = p;
	out->E = E;
	out->F = (int)(R.faces.size() / 6);
	out->xyz = dup(R.xyz);
	out->faces = dup(R.faces);
	out->v_to_v = dup(R.v2v);
	out->v_to_lf = dup(R.v2lf);
	out->v_to_lf_wp = dup(R.v2lfwp);
	out->n_unsupported_bc = R.n_unsupported;
	out->n_inverted = R.n_inverted;
	if (!out->xyz || !out->faces || !out->v_to_v || !out->v_to_lf || !out->v_to_lf_wp) return DGX_ERR_ARG;
	return DGX_OK;
}

// 3-D: first-order HEX volumes (gmsh type 5), QUAD boundary faces (type 3) carrying the physical tags, and
// periodic pairs tagged 51/52 (x), 53/54 (y), 55/56 (z), master first (mesh_periodic.c).  gmsh's HEX vertices
// (bottom face counter-clockwise, then the top face) are taken in tensor-product order (r fastest, then s, t):
// (0, 1, 3, 2, 4, 5, 7, 6).  Local faces 0..5 are r = -1, r = +1, s = -1, s = +1, t = -1, t = +1, each with its
// corners in face order (a, b), a fastest: (s, t) on r faces, (r, t) on s faces, (r, s) on t faces.  A face's
// orientation code is the one of the eight QUAD-face codes whose corner permutation (dgx_face_permutation at
// p = 1) carries the left face's corners onto the right face's.  The geometry is the trilinear map of the
// vertices at the GLL nodes of degree p.
static dgx_status read_hex(const char* path, int p, dgx_host_mesh* out)
{
	std::ifstream in(path);
	if (!in) return DGX_ERR_ARG;
	std::map<int, std::array<double, 3>> nodes;
	std::vector<Elem> elems;
	std::string line;
	while (std::getline(in, line)) {
		if (line.rfind("$Nodes", 0) == 0) {
			int n = 0;
			in >> n;
			for (int i = 0; i < n; ++i) {
				int id = 0;
				double x = 0.0, y = 0.0, z = 0.0;
				if (!(in >> id >> x >> y >> z)) return DGX_ERR_ARG;
				nodes[id] = { x, y, z };
			}
		} else if (line.rfind("$Elements", 0) == 0) {
			int n = 0;
			in >> n;
			std::getline(in, line);
			for (int i = 0; i < n; ++i) {
				std::getline(in, line);
				std::istringstream ss(line);
				int id = 0, type = 0, ntag = -1;
				if (!(ss >> id >> type >> ntag) || ntag < 0 || ntag > 64) return DGX_ERR_ARG;
				Elem e;
				e.type = type;
				e.tags.resize(ntag);
				for (int t = 0; t < ntag; ++t)
					if (!(ss >> e.tags[t])) return DGX_ERR_ARG;
				int v;
				while (ss >> v) e.nodes.push_back(v);
				const int need = type == 15 ? 1 : type == 1 ? 2 : type == 3 ? 4 : type == 5 ? 8 : -1;
				if (need < 0) return DGX_ERR_UNSUPPORTED;   // TET, PRISM, PYR, second order
				if ((int)e.nodes.size() < need) return DGX_ERR_ARG;
				for (int k = 0; k < need; ++k)
					if (!nodes.count(e.nodes[k])) return DGX_ERR_ARG;
				e.nodes.resize(need);
				elems.push_back(e);
			}
		}
	}
	using Key = std::array<int, 4>;
	auto key4 = [](Key k) { std::sort(k.begin(), k.end()); return k; };
	std::vector<std::array<int, 8>> vol;
	std::map<Key, int> btag;
	std::vector<Key> masters[3], slaves[3];
	for (const Elem& e : elems) {
		if (e.type == 5) {
			const std::vector<int>& n = e.nodes;
			vol.push_back({ n[0], n[1], n[3], n[2], n[4], n[5], n[7], n[6] });
		} else if (e.type == 3 && !e.tags.empty()) {
			const Key q = { e.nodes[0], e.nodes[1], e.nodes[2], e.nodes[3] };
			btag[key4(q)] = e.tags[0];
			const int base = e.tags[0] % 10000;
			if (base >= 51 && base <= 56) ((base % 2) ? masters : slaves)[(base - 51) / 2].push_back(q);
			else if (base == 61 || base == 62) return DGX_ERR_UNSUPPORTED;
		}
	}
	const int E = (int)vol.size();
	if (E == 0) return DGX_ERR_UNSUPPORTED;
	// periodic slave vertex -> master vertex, matched on the two coordinates along the face
	double lo[3] = { 1e300, 1e300, 1e300 }, hi[3] = { -1e300, -1e300, -1e300 };
	for (const auto& nd : nodes)
		for (int c = 0; c < 3; ++c) { lo[c] = std::min(lo[c], nd.second[c]); hi[c] = std::max(hi[c], nd.second[c]); }
	const double tol = 1e-9 * std::max({ hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-300 });
	std::map<Key, std::map<int, int>> slave_map;
	for (int dir = 0; dir < 3; ++dir) {
		if (masters[dir].size() != slaves[dir].size()) return DGX_ERR_ARG;
		const int t0 = (dir + 1) % 3, t1 = (dir + 2) % 3;
		auto same = [&](int a, int b) {
			return std::fabs(nodes[a][t0] - nodes[b][t0]) <= tol && std::fabs(nodes[a][t1] - nodes[b][t1]) <= tol;
		};
		for (const Key& S : slaves[dir]) {
			bool found = false;
			for (const Key& Mq : masters[dir]) {
				std::map<int, int> mp;
				for (int a : S)
					for (int b : Mq)
						if (same(a, b)) mp[a] = b;
				if ((int)mp.size() == 4) {
					slave_map[key4(S)] = mp;
					found = true;
					break;
				}
			}
			if (!found) return DGX_ERR_ARG;   // a slave face with no master counterpart
		}
	}
	// local face corners (a + 2 b) as tensor-product vertex indices
	static const int fc[6][4] = { { 0, 2, 4, 6 }, { 1, 3, 5, 7 }, { 0, 1, 4, 5 },
	                              { 2, 3, 6, 7 }, { 0, 1, 2, 3 }, { 4, 5, 6, 7 } };
	const int nf = 6 * E;
	std::vector<Key> raw(nf), cor(nf);
	for (int v = 0; v < E; ++v)
		for (int lf = 0; lf < 6; ++lf) {
			Key r;
			for (int k = 0; k < 4; ++k) r[k] = vol[v][fc[lf][k]];
			raw[6 * v + lf] = r;
			auto it = slave_map.find(key4(r));
			Key c = r;
			if (it != slave_map.end())
				for (int k = 0; k < 4; ++k) c[k] = it->second.at(r[k]);
			cor[6 * v + lf] = c;
		}
	std::map<Key, std::vector<int>> groups;
	for (int f = 0; f < nf; ++f) groups[key4(cor[f])].push_back(f);
	int perms[8][4];
	for (int o = 0; o < 8; ++o)
		if (dgx_face_permutation(3, 1, o, perms[o]) != DGX_OK) return DGX_ERR_ARG;
	std::vector<int> v2v(nf, -1), v2lf(nf, -1), v2lfwp(nf, -1), faces;
	int n_unsupported = 0;
	for (const auto& g : groups) {
		const std::vector<int>& fs = g.second;
		if (fs.size() == 2) {
			const int f0 = std::min(fs[0], fs[1]), f1 = std::max(fs[0], fs[1]);
			v2v[f0] = f1 / 6;
			v2lf[f0] = f1 % 6;
			v2v[f1] = f0 / 6;
			v2lf[f1] = f0 % 6;
			if (key4(raw[f0]) == key4(raw[f1])) {
				v2lfwp[f0] = v2lf[f0];
				v2lfwp[f1] = v2lf[f1];
			}
			int ord = -1;
			for (int o = 0; o < 8 && ord < 0; ++o) {
				bool ok = true;
				for (int k = 0; k < 4 && ok; ++k) ok = cor[f1][perms[o][k]] == cor[f0][k];
				if (ok) ord = o;
			}
			if (ord < 0) return DGX_ERR_ARG;   // the two faces' corners do not coincide
			const int row[6] = { f0 / 6, f0 % 6, f1 / 6, f1 % 6, ord, DGX_BC_INTERIOR };
			faces.insert(faces.end(), row, row + 6);
		} else if (fs.size() == 1) {
			const int f = fs[0];
			auto it = btag.find(key4(raw[f]));
			const int tag = (it == btag.end()) ? -1 : it->second;
			v2lf[f] = tag;
			const int bc = bc_code(tag);
			if (bc < 0) ++n_unsupported;
			const int row[6] = { f / 6, f % 6, -1, -1, -1, bc < 0 ? DGX_BC_RIEMANN : bc };
			faces.insert(faces.end(), row, row + 6);
		} else {
			return DGX_ERR_UNSUPPORTED;   // non-manifold face
		}
	}
	for (int f = 0; f < nf; ++f)
		if (v2lfwp[f] == -1) {
			auto it = btag.find(key4(raw[f]));
			if (it != btag.end()) v2lfwp[f] = it->second;
		}
	const int N = p + 1, Np = N * N * N;
	std::vector<double> r, xyz((size_t)E * 3 * Np);
	gll_nodes(N, r);
	int n_inverted = 0;
	for (int v = 0; v < E; ++v) {
		double X[8][3];
		for (int k = 0; k < 8; ++k)
			for (int c = 0; c < 3; ++c) X[k][c] = nodes[vol[v][k]][c];
		for (int k2 = 0; k2 < N; ++k2)
			for (int j2 = 0; j2 < N; ++j2)
				for (int i2 = 0; i2 < N; ++i2) {
					const double w[2][3] = { { 0.5 * (1 - r[i2]), 0.5 * (1 - r[j2]), 0.5 * (1 - r[k2]) },
					                         { 0.5 * (1 + r[i2]), 0.5 * (1 + r[j2]), 0.5 * (1 + r[k2]) } };
					for (int c = 0; c < 3; ++c) {
						double x = 0.0;
						for (int k = 0; k < 8; ++k) x += w[k & 1][0] * w[(k >> 1) & 1][1] * w[k >> 2][2] * X[k][c];
						xyz[((size_t)v * 3 + c) * Np + i2 + N * (j2 + N * k2)] = x;
					}
				}
		double J[3][3];
		for (int c = 0; c < 3; ++c) {
			J[c][0] = X[1][c] - X[0][c];
			J[c][1] = X[2][c] - X[0][c];
			J[c][2] = X[4][c] - X[0][c];
		}
		const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
		                   J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
		                   J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
		if (det <= 0.0) ++n_inverted;
	}
	out->d = 3;
	out->p = p;
	out->E = E;
	out->F = (int)(faces.size() / 6);
	out->xyz = dup(xyz);
	out->faces = dup(faces);
	out->v_to_v = dup(v2v);
	out->v_to_lf = dup(v2lf);
	out->v_to_lf_wp = dup(v2lfwp);
	out->n_unsupported_bc = n_unsupported;
	out->n_inverted = n_inverted;
	if (!out->xyz || !out->faces || !out->v_to_v || !out->v_to_lf || !out->v_to_lf_wp) return DGX_ERR_ARG;
	return DGX_OK;
}

static dgx_status read_gmsh_mixed(const char* path, int p, dgx_host_mesh_mixed* out)
{
	ReadOut R;
	const dgx_status st = read_any(path, p, true, R);
	if (st != DGX_OK) return st;
	out->E = (int)R.type.size();
	out->F = (int)(R.faces.size() / 6);
	out->p = p;
	out->type = dup(R.type);
	out->xyz = dup(R.xyz);
	out->faces = dup(R.faces);
	out->v_to_v = dup(R.v2v);
	out->v_to_lf = dup(R.v2lf);
	out->v_to_lf_wp = dup(R.v2lfwp);
	out->n_unsupported_bc = R.n_unsupported;
	out->n_inverted = R.n_inverted;
	if (!out->type || !out->xyz || !out->faces || !out->v_to_v || !out->v_to_lf || !out->v_to_lf_wp) return DGX_ERR_ARG;
	return DGX_OK;
}

extern "C" {

void dgx_mesh_free(dgx_host_mesh* m)
{
	if (!m) return;
	std::free(m->xyz);
	std::free(m->faces);
	std::free(m->v_to_v);
	std::free(m->v_to_lf);
	std::free(m->v_to_lf_wp);
	std::memset(m, 0, sizeof(*m));
}

dgx_status dgx_mesh_read_gmsh(const char* path, int p, dgx_host_mesh* out)
{
	if (!path || !out || p < 1 || p > 5) return DGX_ERR_ARG;
	std::memset(out, 0, sizeof(*out));
	dgx_status st;
	try {
		st = read_gmsh(path, p, out);
		if (st == DGX_ERR_UNSUPPORTED) {   // not a 2-D QUAD mesh: a HEX mesh?
			dgx_mesh_free(out);
			std::memset(out, 0, sizeof(*out));
			st = read_hex(path, p, out);
		}
	} catch (...) {   // a malformed file (std::out_of_range, std::bad_alloc): a status code, never an abort
		st = DGX_ERR_ARG;
	}
	if (st != DGX_OK) dgx_mesh_free(out);
	return st;
}

void dgx_mesh_mixed_free(dgx_host_mesh_mixed* m)
{
	if (!m) return;
	std::free(m->type);
	std::free(m->xyz);
	std::free(m->faces);
	std::free(m->v_to_v);
	std::free(m->v_to_lf);
	std::free(m->v_to_lf_wp);
	std::memset(m, 0, sizeof(*m));
}

dgx_status dgx_mesh_read_gmsh_mixed(const char* path, int p, dgx_host_mesh_mixed* out)
{
	if (!path || !out || p < 1 || p > 5) return DGX_ERR_ARG;
	std::memset(out, 0, sizeof(*out));
	dgx_status st;
	try {
		st = read_gmsh_mixed(path, p, out);
	} catch (...) {
		st = DGX_ERR_ARG;
	}
	if (st != DGX_OK) dgx_mesh_mixed_free(out);
	return st;
}

}  // extern "C"
