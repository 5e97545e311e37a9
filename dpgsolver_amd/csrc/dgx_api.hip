// dgx_api.hip -- the C ABI of libdgx (declared in include/dgx.h).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <algorithm>
#include <cstddef>
#include <cstring>
#include <new>
#include <vector>

#include "dgx_internal.h"

using dgx::Ctx;

namespace {

dgx_status hip_status(hipError_t e) { return e == hipSuccess ? DGX_OK : DGX_ERR_HIP; }

#define DGX_TRY(expr)                                  \
	do {                                               \
		hipError_t e_ = (expr);                        \
		if (e_ != hipSuccess) return DGX_ERR_HIP;      \
	} while (0)

bool supported(int d, int p, int pde)
{
	if (d == 2) return p >= 1 && p <= 5;   // Navier-Stokes kernels: N = 2..6 as well
	if (d == 3) return pde == DGX_PDE_EULER && p >= 1 && p <= 4;
	return false;
}

template <typename T>
void release(T*& p)
{
	if (p) (void)hipFree((void*)p);
	p = nullptr;
}

// a device copy of n doubles rounded to FP32
hipError_t upload_f32(float** dst, const double* src, size_t n)
{
	std::vector<float> h(src, src + n);
	hipError_t e = hipMalloc(dst, sizeof(float) * std::max<size_t>(n, 1));
	if (e == hipSuccess && n) e = hipMemcpy(*dst, h.data(), sizeof(float) * n, hipMemcpyHostToDevice);
	return e;
}

void free_hp(Ctx& c)
{
	if (!c.hp) return;
	dgx::HpData* H = c.hp;
	for (dgx_ctx* g : H->groups) dgx_destroy(g);
	release(H->faces); release(H->nJ); release(H->nJ32); release(H->nJhi); release(H->z64); release(H->z32);
	release(H->ops64); release(H->ops32);
	for (dgx::TriGroup& t : H->tri) {
		release(t.met); release(t.detJ); release(t.tr); release(t.jinv);
		release(t.minv); release(t.K); release(t.gv); release(t.gvt);
		release(t.met32); release(t.jinv32); release(t.minv32); release(t.K32);
	}
	release(H->tri_ops64); release(H->tri_ops32); release(H->mfaces); release(H->mnJ); release(H->mnJ32);
	release(H->mxy);
	for (int k = 0; k < dgx::HpData::kLanes; ++k) {
		if (H->lane[k]) (void)hipStreamDestroy(H->lane[k]);
		if (H->join_ev[k]) (void)hipEventDestroy(H->join_ev[k]);
	}
	if (H->fork_ev) (void)hipEventDestroy(H->fork_ev);
	delete H;
	c.hp = nullptr;
}

void free_mesh(Ctx& c)
{
	free_hp(c);
	release(c.xyz); release(c.faces); release(c.elem_faces); release(c.order); release(c.hmin); release(c.metrics_vc); release(c.metrics_vc32);
	release(c.detJ_vc); release(c.detJ_vc32); release(c.nJ_fc); release(c.nJ_fc32); release(c.u0); release(c.rtmp);
	release(c.color_elems);
	release(c.ns_gvt); release(c.ns_kappa); release(c.wall_xy); release(c.ns_gvt_dual); release(c.ns_kappa_dual);
	c.color_off.clear();
	for (int& g : c.rhs_grid) g = 0;   // launch shapes are tuned per mesh
	c.mesh_ready = false;
}

size_t ndof(const Ctx& c) { return c.hp ? (size_t)c.hp->ndof : (size_t)c.E * c.Np * c.nv; }

dgx_status upload_uniform(dgx_ctx* h, const dgx_mesh* m, bool nonconf_ok);
void lagrange_at(const double* xn, int n, double x, double* l, double* dl);
int face_node_2d(int N, int lf, int i);

}  // namespace

extern "C" {

const char* dgx_version(void) { return "dgx 0.1 (gfx950)"; }

dgx_status dgx_device_count(int* n)
{
	if (!n) return DGX_ERR_ARG;
	return hip_status(hipGetDeviceCount(n));
}

dgx_status dgx_create(dgx_ctx** out, const dgx_config* cfg)
{
	if (!out || !cfg) return DGX_ERR_ARG;
	*out = nullptr;
	if (cfg->pde != DGX_PDE_EULER && cfg->pde != DGX_PDE_ADVECTION && cfg->pde != DGX_PDE_NAVIER_STOKES)
		return DGX_ERR_ARG;
	if (cfg->pde != DGX_PDE_ADVECTION && cfg->num_flux != DGX_FLUX_LF && cfg->num_flux != DGX_FLUX_ROE) return DGX_ERR_ARG;
	if (cfg->pde == DGX_PDE_NAVIER_STOKES) {
		if (cfg->d != 2) return DGX_ERR_UNSUPPORTED;
		if (!(cfg->mu >= 0.0) || !(cfg->Pr > 0.0)) return DGX_ERR_ARG;
		if (cfg->viscosity_type != DGX_VISC_CONSTANT && cfg->viscosity_type != DGX_VISC_SUTHERLAND) return DGX_ERR_ARG;
		if (cfg->viscosity_type == DGX_VISC_SUTHERLAND && !(cfg->r_s >= 0.0)) return DGX_ERR_ARG;
	}
	if (cfg->pde == DGX_PDE_ADVECTION && cfg->num_flux != DGX_FLUX_UPWIND) return DGX_ERR_ARG;
	if (!supported(cfg->d, cfg->p, cfg->pde)) return DGX_ERR_UNSUPPORTED;
	dgx_ctx* h = new (std::nothrow) dgx_ctx();
	if (!h) return DGX_ERR_ARG;
	Ctx& c = h->c;
	c.cfg = *cfg;
	c.d = cfg->d;
	c.p = cfg->p;
	c.N = cfg->p + 1;
	c.Np = c.N * c.N * (c.d == 3 ? c.N : 1);
	c.Nc = c.Np;
	c.Nfc = (c.d == 3) ? c.N * c.N : c.N;
	c.nv = (cfg->pde == DGX_PDE_ADVECTION) ? 1 : c.d + 2;
	if (hipGetDevice(&c.device) != hipSuccess) { delete h; return DGX_ERR_HIP; }
	// the no-allocation entry point's static scratch array on this device (looked up once per device; a
	// failure here only disables dgx_ssp_rk3_step_noalloc)
	if (dgx::noalloc_scratch(c.device, &c.noalloc_scratch) != hipSuccess) c.noalloc_scratch = nullptr;
	dgx::build_ops(c.d, c.p, c.h_ops64, c.h_ops32);
	if (hipMalloc(&c.ops64, sizeof(dgx::Ops1D<double>)) != hipSuccess ||
	    hipMalloc(&c.ops32, sizeof(dgx::Ops1D<float>)) != hipSuccess ||
	    hipMalloc(&c.dmax, sizeof(unsigned long long) * 4) != hipSuccess ||
	    hipMalloc(&c.dstatus, sizeof(int) * 4) != hipSuccess) {
		dgx_destroy(h);
		return DGX_ERR_HIP;
	}
	if (hipMemcpy(c.ops64, &c.h_ops64, sizeof(c.h_ops64), hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(c.ops32, &c.h_ops32, sizeof(c.h_ops32), hipMemcpyHostToDevice) != hipSuccess) {
		dgx_destroy(h);
		return DGX_ERR_HIP;
	}
	*out = h;
	return DGX_OK;
}

dgx_status dgx_destroy(dgx_ctx* h)
{
	if (!h) return DGX_OK;
	Ctx& c = h->c;
	free_mesh(c);
	release(c.ops64); release(c.ops32); release(c.dmax); release(c.dstatus);
	release(c.blocks); release(c.lu); release(c.piv);
	delete h;
	return DGX_OK;
}

dgx_status dgx_set_stream(dgx_ctx* h, void* stream)
{
	if (!h) return DGX_ERR_ARG;
	h->c.stream = (hipStream_t)stream;
	return DGX_OK;
}

dgx_status dgx_synchronize(dgx_ctx* h)
{
	if (!h) return DGX_ERR_ARG;
	return hip_status(hipStreamSynchronize(h->c.stream));
}

dgx_status dgx_sizes(const dgx_ctx* h, int* Np, int* nv, int* Nfc, long long* nd)
{
	if (!h) return DGX_ERR_ARG;
	if (Np) *Np = h->c.Np;
	if (nv) *nv = h->c.nv;
	if (Nfc) *Nfc = h->c.Nfc;
	if (nd) *nd = (long long)ndof(h->c);
	return DGX_OK;
}

dgx_status dgx_upload_mesh(dgx_ctx* h, const dgx_mesh* m) { return upload_uniform(h, m, false); }

}  // extern "C"

namespace {

// The uniform-order upload; nonconf_ok admits faces marked kBcNonconforming (the order groups of a
// variable-order mesh).
dgx_status upload_uniform(dgx_ctx* h, const dgx_mesh* m, bool nonconf_ok)
{
	if (!h || !m || m->E <= 0 || m->F < 0 || !m->xyz || (m->F > 0 && !m->faces)) return DGX_ERR_ARG;
	Ctx& c = h->c;
	const int d = c.d, NFACE = 2 * d;
	// validate the face table and build the per-element face records
	std::vector<dgx::ElemFace> ef((size_t)m->E * NFACE, dgx::ElemFace{ -1, -1, -1, 0 });
	for (int f = 0; f < m->F; ++f) {
		const int* r = m->faces + 6 * f;
		const int eL = r[0], fL = r[1], eR = r[2], fR = r[3], bc = r[5];
		if (eL < 0 || eL >= m->E || fL < 0 || fL >= NFACE) return DGX_ERR_ARG;
		// the reference's ind_ord -> permutation bits; boundary faces carry none (the reference writes -1)
		const int ord = (bc == DGX_BC_INTERIOR) ? dgx::ord_bits(d, r[4]) : 0;
		if (ord < 0 || (bc != DGX_BC_INTERIOR && r[4] != -1 && dgx::ord_bits(d, r[4]) < 0)) return DGX_ERR_ARG;
		if (bc == DGX_BC_INTERIOR) {
			if (eR < 0 || eR >= m->E || fR < 0 || fR >= NFACE) return DGX_ERR_ARG;
		} else if (nonconf_ok && bc == dgx::kBcNonconforming) {
			// face to an element of another order: its flux arrives lifted from the face kernel
		} else {
			// Euler: Riemann and slip walls; Navier-Stokes (2-D): Riemann and the no-slip walls
			// (set_function_pointers_num_flux_bc_{euler,navier_stokes}, face_solver_T.c:251-298)
			const bool noslip = bc == DGX_BC_NOSLIP_ADIABATIC || bc == DGX_BC_NOSLIP_DIABATIC || bc == DGX_BC_NOSLIP_ROTATING;
			if (bc != DGX_BC_RIEMANN && bc != DGX_BC_SLIPWALL && !noslip) return DGX_ERR_ARG;
			if (c.cfg.pde == DGX_PDE_NAVIER_STOKES) {
				if (d != 2 || bc == DGX_BC_SLIPWALL) return DGX_ERR_UNSUPPORTED;
			} else if (c.cfg.pde != DGX_PDE_EULER) {
				return DGX_ERR_ARG;
			} else if (noslip) {
				return DGX_ERR_UNSUPPORTED;
			}
		}
		dgx::ElemFace& L = ef[(size_t)eL * NFACE + fL];
		if (L.f >= 0) return DGX_ERR_ARG;  // local face used twice
		L = dgx::ElemFace{ f, bc == DGX_BC_INTERIOR ? eR : -1, (bc == DGX_BC_INTERIOR ? fR : 0) | (0 << 4) | (ord << 8) | (bc << 12), 0 };
		if (bc == DGX_BC_INTERIOR) {
			dgx::ElemFace& R = ef[(size_t)eR * NFACE + fR];
			if (R.f >= 0) return DGX_ERR_ARG;
			R = dgx::ElemFace{ f, eL, fL | (1 << 4) | (ord << 8), 0 };
		}
	}
	bool nonconf = false;
	for (const auto& x : ef) {
		if (x.f < 0) return DGX_ERR_ARG;  // open mesh: every local face needs a face record
		if (((x.packed >> 12) & 15) == dgx::kBcNonconforming) nonconf = true;
	}
	free_mesh(c);
	c.has_nonconforming = nonconf;
	c.E = m->E;
	c.F = m->F;
	const size_t nx = (size_t)c.E * d * c.Np, nm = (size_t)c.E * d * d * c.Nc, nj = (size_t)c.E * c.Nc,
	             nn = (size_t)(c.F > 0 ? c.F : 1) * c.Nfc * d;
	DGX_TRY(hipMalloc(&c.xyz, sizeof(double) * nx));
	DGX_TRY(hipMalloc(&c.faces, sizeof(int) * 6 * (size_t)(c.F > 0 ? c.F : 1)));
	DGX_TRY(hipMalloc(&c.elem_faces, sizeof(dgx::ElemFace) * ef.size()));
	DGX_TRY(hipMalloc(&c.metrics_vc, sizeof(double) * nm));
	DGX_TRY(hipMalloc(&c.metrics_vc32, sizeof(float) * nm));
	DGX_TRY(hipMalloc(&c.detJ_vc, sizeof(double) * nj));
	DGX_TRY(hipMalloc(&c.detJ_vc32, sizeof(float) * nj));
	DGX_TRY(hipMalloc(&c.nJ_fc, sizeof(double) * nn));
	DGX_TRY(hipMalloc(&c.nJ_fc32, sizeof(float) * nn));
	DGX_TRY(hipMemcpy(c.xyz, m->xyz, sizeof(double) * nx, hipMemcpyHostToDevice));
	if (c.F > 0) DGX_TRY(hipMemcpy(c.faces, m->faces, sizeof(int) * 6 * (size_t)c.F, hipMemcpyHostToDevice));
	DGX_TRY(hipMemcpy(c.elem_faces, ef.data(), sizeof(dgx::ElemFace) * ef.size(), hipMemcpyHostToDevice));
	{
		const std::vector<int> order = dgx::morton_order(d, c.E, c.Np, m->xyz);
		DGX_TRY(hipMalloc(&c.order, sizeof(int) * order.size()));
		DGX_TRY(hipMemcpy(c.order, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice));
		const std::vector<double> hm = dgx::min_edge_lengths(d, c.N, c.E, m->xyz);
		DGX_TRY(hipMalloc(&c.hmin, sizeof(double) * hm.size()));
		DGX_TRY(hipMemcpy(c.hmin, hm.data(), sizeof(double) * hm.size(), hipMemcpyHostToDevice));
		// greedy colouring of the face-neighbour graph (multicolour block Gauss-Seidel)
		std::vector<int> color(c.E, -1);
		int ncol = 0;
		for (int e = 0; e < c.E; ++e) {
			unsigned used = 0;
			for (int lf = 0; lf < NFACE; ++lf) {
				const int n = ef[(size_t)e * NFACE + lf].en;
				if (n >= 0 && n != e && color[n] >= 0 && color[n] < 32) used |= 1u << color[n];
			}
			int k = 0;
			while (k < 32 && (used >> k) & 1u) ++k;
			color[e] = k;
			ncol = std::max(ncol, k + 1);
		}
		c.color_off.assign(ncol + 1, 0);
		for (int e = 0; e < c.E; ++e) c.color_off[color[e] + 1] += 1;
		for (int k = 0; k < ncol; ++k) c.color_off[k + 1] += c.color_off[k];
		std::vector<int> elems(c.E), fill(c.color_off.begin(), c.color_off.end() - 1);
		for (int e = 0; e < c.E; ++e) elems[fill[color[e]]++] = e;
		DGX_TRY(hipMalloc(&c.color_elems, sizeof(int) * c.E));
		DGX_TRY(hipMemcpy(c.color_elems, elems.data(), sizeof(int) * c.E, hipMemcpyHostToDevice));
	}
	if (c.cfg.pde == DGX_PDE_NAVIER_STOKES) {
		// face traces of the volume gradient and the BR2 lift trace factors
		DGX_TRY(hipMalloc(&c.ns_gvt, sizeof(double) * (size_t)c.E * NFACE * d * c.nv * c.Nfc));
		DGX_TRY(hipMalloc(&c.ns_kappa, sizeof(double) * (size_t)c.E * NFACE * c.Nfc));
		// rotating walls: the physical coordinates of the face GL points (the wall velocity omega (-y, x) there),
		// interpolated from the element's GLL nodes on the face
		bool rotating = false;
		for (int f = 0; f < m->F; ++f) rotating |= m->faces[6 * f + 5] == DGX_BC_NOSLIP_ROTATING;
		if (rotating) {
			const int N = c.N, Np = c.Np;
			double xl[dgx::kMaxN], wl[dgx::kMaxN], xg[dgx::kMaxN], wg[dgx::kMaxN], l[dgx::kMaxN], dl[dgx::kMaxN];
			dgx::gauss_lobatto(N, xl, wl);
			dgx::gauss_legendre(N, xg, wg);
			std::vector<double> xy((size_t)c.F * N * 2, 0.0);
			for (int f = 0; f < m->F; ++f) {
				const int* r = m->faces + 6 * f;
				if (r[5] != DGX_BC_NOSLIP_ROTATING) continue;
				const double* xe = m->xyz + (size_t)r[0] * 2 * Np;
				for (int q = 0; q < N; ++q) {
					lagrange_at(xl, N, xg[q], l, dl);
					double px = 0.0, py = 0.0;
					for (int i = 0; i < N; ++i) {
						const int vn = face_node_2d(N, r[1], i);
						px += l[i] * xe[vn];
						py += l[i] * xe[Np + vn];
					}
					xy[((size_t)f * N + q) * 2] = px;
					xy[((size_t)f * N + q) * 2 + 1] = py;
				}
			}
			DGX_TRY(hipMalloc(&c.wall_xy, sizeof(double) * xy.size()));
			DGX_TRY(hipMemcpy(c.wall_xy, xy.data(), sizeof(double) * xy.size(), hipMemcpyHostToDevice));
		}
	}
	DGX_TRY(dgx::launch_geometry(c));
	DGX_TRY(dgx::launch_convert(c.metrics_vc, c.metrics_vc32, nm, c.stream));
	DGX_TRY(dgx::launch_convert(c.detJ_vc, c.detJ_vc32, nj, c.stream));
	DGX_TRY(dgx::launch_convert(c.nJ_fc, c.nJ_fc32, nn, c.stream));
	DGX_TRY(hipStreamSynchronize(c.stream));
	c.mesh_ready = true;
	return DGX_OK;
}

// Lagrange basis of the nodes xn (n) and its derivative at x.
void lagrange_at(const double* xn, int n, double x, double* l, double* dl)
{
	for (int i = 0; i < n; ++i) {
		double li = 1.0, dli = 0.0;
		for (int j = 0; j < n; ++j) {
			if (j == i) continue;
			const double r = 1.0 / (xn[i] - xn[j]);
			dli = dli * (x - xn[j]) * r + li * r;
			li *= (x - xn[j]) * r;
		}
		l[i] = li;
		dl[i] = dli;
	}
}

void build_hp_ops(dgx::HpOps<double>& o)
{
	std::memset(&o, 0, sizeof(o));
	double xl[dgx::kHpMaxN + 1][dgx::kHpMaxN], xg[dgx::kHpMaxN + 1][dgx::kHpMaxN], wg[dgx::kHpMaxN + 1][dgx::kHpMaxN], w[dgx::kHpMaxN];
	for (int N = 2; N <= dgx::kHpMaxN; ++N) {
		dgx::gauss_lobatto(N, xl[N], w);
		dgx::gauss_legendre(N, xg[N], wg[N]);
		for (int i = 0; i < N; ++i) {
			o.w[N][i] = wg[N][i];
			o.winv[N][i] = 1.0 / wg[N][i];
		}
	}
	double l[dgx::kHpMaxN], dl[dgx::kHpMaxN];
	for (int Ns = 2; Ns <= dgx::kHpMaxN; ++Ns)
		for (int Nf = 2; Nf <= dgx::kHpMaxN; ++Nf)
			for (int q = 0; q < Nf; ++q) {
				lagrange_at(xl[Ns], Ns, xg[Nf][q], l, dl);
				for (int i = 0; i < Ns; ++i) o.Ig[Ns][Nf][q * Ns + i] = l[i];
				lagrange_at(xg[Ns], Ns, xg[Nf][q], l, dl);
				for (int i = 0; i < Ns; ++i) o.Pg[Ns][Nf][q * Ns + i] = l[i];
			}
}

int face_node_2d(int N, int lf, int i)
{
	const int fix = (lf & 1) ? N - 1 : 0;
	return (lf >> 1) == 0 ? fix + N * i : i + N * fix;
}

// n J_f (outward) of local face lf of an element of N GLL nodes per direction at the face coordinates s
// (face_solver.c:95-107 takes the face geometry from one volume): the rotated tangent of the face
// interpolant, J grad xi = (y_eta, -x_eta) on xi faces, J grad eta = (-y_xi, x_xi) on eta faces.
void outward_nj(const double* xyz, int N, int lf, const double* s, int ns, double* nj)
{
	double xl[dgx::kHpMaxN], w[dgx::kHpMaxN], l[dgx::kHpMaxN], dl[dgx::kHpMaxN];
	dgx::gauss_lobatto(N, xl, w);
	const int Np = N * N;
	for (int q = 0; q < ns; ++q) {
		lagrange_at(xl, N, s[q], l, dl);
		double tx = 0.0, ty = 0.0;
		for (int i = 0; i < N; ++i) {
			const int vn = face_node_2d(N, lf, i);
			tx += dl[i] * xyz[vn];
			ty += dl[i] * xyz[Np + vn];
		}
		double nx, ny;
		if ((lf >> 1) == 0) { nx = ty; ny = -tx; }
		else { nx = -ty; ny = tx; }
		const double sg = (lf & 1) ? 1.0 : -1.0;
		nj[2 * q] = sg * nx;
		nj[2 * q + 1] = sg * ny;
	}
}

// Enqueues launch(g, s) for every order group on the group's side stream.  Each group context gets its own
// stream back afterwards (not the caller's current one: during a graph capture that is the private capture
// stream, destroyed when the capture ends, and a group context must never keep it).
template <typename F>
hipError_t on_groups(Ctx& c, F launch)
{
	dgx::HpData& H = *c.hp;
	for (size_t g = 0; g < H.groups.size(); ++g) {
		Ctx& s = H.groups[g]->c;
		const hipStream_t own = s.stream;
		s.stream = H.lane[H.group_lane[g]];
		const hipError_t e = launch(g, s);
		s.stream = own;
		if (e != hipSuccess) return e;
	}
	return hipSuccess;
}

// Every side stream waits for everything enqueued so far on the side streams.
hipError_t hp_barrier(dgx::HpData& H, hipStream_t main)
{
	const hipError_t e = dgx::hp_join(H, main);
	return e != hipSuccess ? e : dgx::hp_fork(H, main);
}

// Variable-order and mixed-type residuals (FP64). Navier-Stokes: every group's gradient traces (TRI: and state
// traces), the faces between orders or types (lifted gradients and fluxes; one launch per face-point count),
// every group's residual (+ M^-1 and max-norm). Euler: the TRI state traces, the faces, every group's residual.
// The launches of one phase write disjoint outputs and run concurrently on the side streams; the phases are
// ordered by barriers.
dgx_status hp_rhs64(Ctx& c, const double* u, double* r, int mode, unsigned long long* dmax)
{
	dgx::HpData& H = *c.hp;
	DGX_TRY(dgx::hp_fork(H, c.stream));
	if (!H.tri.empty() && mode != dgx::MODE_RHS && mode != dgx::MODE_RHS_MINV) {
		(void)dgx::hp_join(H, c.stream);
		return DGX_ERR_UNSUPPORTED;
	}
	const bool minv = (mode == dgx::MODE_RHS_MINV);
	if (c.cfg.pde == DGX_PDE_NAVIER_STOKES) {
		const double mu = c.cfg.mu, Pr = c.cfg.Pr;
		DGX_TRY(on_groups(c, [&](size_t g, Ctx& s) { return dgx::launch_rhs_ns_phase(s, u + H.u_off[g], r + H.u_off[g], 1); }));
		for (size_t t = 0; t < H.tri.size(); ++t)
			DGX_TRY(dgx::launch_tri_ns_phase64(H.tri[t], H.tri_ops64, u, r, H.z64, mu, Pr, 1, H.lane[H.tri_lane[t]]));
		DGX_TRY(hp_barrier(H, c.stream));
		if (H.tri.empty()) DGX_TRY(dgx::launch_hp_faces64(c, u));
		else DGX_TRY(dgx::launch_mix_faces_ns64(H, c, u, H.lane[dgx::HpData::kLanes - 1]));
		DGX_TRY(hp_barrier(H, c.stream));
		for (size_t t = 0; t < H.tri.size(); ++t) {
			DGX_TRY(dgx::launch_tri_ns_phase64(H.tri[t], H.tri_ops64, u, r, H.z64, mu, Pr, 2, H.lane[H.tri_lane[t]]));
			if (minv) DGX_TRY(dgx::launch_tri_minv64(H.tri[t], H.tri_ops64, r, dmax, H.lane[H.tri_lane[t]]));
		}
		DGX_TRY(on_groups(c, [&](size_t g, Ctx& s) -> hipError_t {
			double* rg = r + H.u_off[g];
			hipError_t e = dgx::launch_rhs_ns_phase(s, u + H.u_off[g], rg, 2);
			if (e != hipSuccess || mode != dgx::MODE_RHS_MINV) return e;
			e = dgx::launch_minv(s, rg);
			if (e == hipSuccess && dmax) e = dgx::launch_maxabs(s, rg, ndof(s), dmax);
			return e;
		}));
	} else {
		if (H.tri.empty()) {
			DGX_TRY(dgx::launch_hp_faces64(c, u));
		} else {   // the TRI traces, then the faces
			for (size_t t = 0; t < H.tri.size(); ++t)
				DGX_TRY(dgx::launch_tri_trace64(H.tri[t], H.tri_ops64, u, H.lane[H.tri_lane[t]]));
			DGX_TRY(hp_barrier(H, c.stream));
			DGX_TRY(dgx::launch_mix_faces64(H, c, u, H.lane[dgx::HpData::kLanes - 1]));
		}
		DGX_TRY(hp_barrier(H, c.stream));
		DGX_TRY(on_groups(c, [&](size_t g, Ctx& s) { return dgx::launch_rhs64(s, u + H.u_off[g], r + H.u_off[g], mode, dmax); }));
		for (size_t t = 0; t < H.tri.size(); ++t) {
			DGX_TRY(dgx::launch_tri_rhs64(H.tri[t], H.tri_ops64, u, r, H.z64, H.lane[H.tri_lane[t]]));
			if (minv) DGX_TRY(dgx::launch_tri_minv64(H.tri[t], H.tri_ops64, r, dmax, H.lane[H.tri_lane[t]]));
		}
	}
	DGX_TRY(dgx::hp_join(H, c.stream));
	return DGX_OK;
}

dgx_status hp_rhs32(Ctx& c, const float* u, float* r)
{
	dgx::HpData& H = *c.hp;
	DGX_TRY(dgx::hp_fork(H, c.stream));
	if (c.cfg.pde == DGX_PDE_NAVIER_STOKES) {
		const double mu = c.cfg.mu, Pr = c.cfg.Pr;
		DGX_TRY(on_groups(c, [&](size_t g, Ctx& s) { return dgx::launch_rhs_ns32_phase(s, u + H.u_off[g], r + H.u_off[g], 1); }));
		for (size_t t = 0; t < H.tri.size(); ++t)
			DGX_TRY(dgx::launch_tri_ns_phase32(H.tri[t], H.tri_ops32, u, r, H.z32, mu, Pr, 1, H.lane[H.tri_lane[t]]));
		DGX_TRY(hp_barrier(H, c.stream));
		if (H.tri.empty()) DGX_TRY(dgx::launch_hp_faces32(c, u));
		else DGX_TRY(dgx::launch_mix_faces32(H, c, u, true));
		DGX_TRY(hp_barrier(H, c.stream));
		for (size_t t = 0; t < H.tri.size(); ++t)
			DGX_TRY(dgx::launch_tri_ns_phase32(H.tri[t], H.tri_ops32, u, r, H.z32, mu, Pr, 2, H.lane[H.tri_lane[t]]));
		DGX_TRY(on_groups(c, [&](size_t g, Ctx& s) { return dgx::launch_rhs_ns32_phase(s, u + H.u_off[g], r + H.u_off[g], 2); }));
	} else {
		if (H.tri.empty()) {
			DGX_TRY(dgx::launch_hp_faces32(c, u));
		} else {   // the TRI traces, then the faces
			for (size_t t = 0; t < H.tri.size(); ++t)
				DGX_TRY(dgx::launch_tri_trace32(H.tri[t], H.tri_ops32, u, H.lane[H.tri_lane[t]]));
			DGX_TRY(hp_barrier(H, c.stream));
			DGX_TRY(dgx::launch_mix_faces32(H, c, u, false));
		}
		DGX_TRY(hp_barrier(H, c.stream));
		DGX_TRY(on_groups(c, [&](size_t g, Ctx& s) { return dgx::launch_rhs32(s, u + H.u_off[g], r + H.u_off[g]); }));
		for (size_t t = 0; t < H.tri.size(); ++t)
			DGX_TRY(dgx::launch_tri_rhs32(H.tri[t], H.tri_ops32, u, r, H.z32, H.lane[H.tri_lane[t]]));
	}
	DGX_TRY(dgx::hp_join(H, c.stream));
	return DGX_OK;
}

// Mixed form: FP64 state in, FP32 residual; the faces between orders in FP64, the groups' mixed kernels.
dgx_status hp_rhs_mixed(Ctx& c, const double* u, float* r)
{
	dgx::HpData& H = *c.hp;
	if (!H.tri.empty()) return DGX_ERR_UNSUPPORTED;   // TRI elements: FP64 only
	DGX_TRY(dgx::hp_fork(H, c.stream));
	if (c.cfg.pde == DGX_PDE_NAVIER_STOKES) {
		DGX_TRY(on_groups(c, [&](size_t g, Ctx& s) { return dgx::launch_rhs_ns_mixed_phase(s, u + H.u_off[g], r + H.u_off[g], 1); }));
		DGX_TRY(hp_barrier(H, c.stream));
		DGX_TRY(dgx::launch_hp_faces_mixed(c, u));
		DGX_TRY(hp_barrier(H, c.stream));
		DGX_TRY(on_groups(c, [&](size_t g, Ctx& s) { return dgx::launch_rhs_ns_mixed_phase(s, u + H.u_off[g], r + H.u_off[g], 2); }));
	} else {
		DGX_TRY(dgx::launch_hp_faces_mixed(c, u));
		DGX_TRY(hp_barrier(H, c.stream));
		DGX_TRY(on_groups(c, [&](size_t g, Ctx& s) { return dgx::launch_rhs_mixed(s, u + H.u_off[g], r + H.u_off[g]); }));
	}
	DGX_TRY(dgx::hp_join(H, c.stream));
	return DGX_OK;
}

}  // namespace

extern "C" {

// Variable-order 2-D mesh of QUAD (type 0) and, when type is given, TRI (type 1) elements sorted by (type, p).
static dgx_status upload_hp_impl(dgx_ctx* h, int E, int F, const int* type, const int* pe, const double* xyz,
                                 const int* faces)
{
	if (!h || E <= 0 || F < 0 || !pe || !xyz || (F > 0 && !faces)) return DGX_ERR_ARG;
	Ctx& c = h->c;
	if (c.d != 2) return DGX_ERR_UNSUPPORTED;
	if (c.cfg.pde != DGX_PDE_EULER && c.cfg.pde != DGX_PDE_NAVIER_STOKES) return DGX_ERR_UNSUPPORTED;
	const int nv = c.nv;
	// per element face: the lifted flux [v][a] (Navier-Stokes: then the projected BR2 products [d][v][a])
	const int zblk = (c.cfg.pde == DGX_PDE_NAVIER_STOKES) ? 3 * nv : nv;
	auto kind = [&](int e) { return type ? type[e] : 0; };
	// QUAD groups of equal order and TRI groups of equal order (the elements arrive sorted by (type, p))
	std::vector<int> grp(E), order, e_off;   // QUAD groups: index into order / e_off; TRI elements: TRI group
	std::vector<int> tri_p, tri_e0;
	bool any_tri = false;
	for (int e = 0; e < E; ++e) {
		const int p = pe[e], k = kind(e);
		if (k != 0 && k != 1) return DGX_ERR_ARG;
		if (p < 1 || p > c.p || p + 1 > dgx::kHpMaxN || (k == 1 && p > dgx::kTriMaxP)) return DGX_ERR_ARG;
		if (e > 0 && k * 8 + p < kind(e - 1) * 8 + pe[e - 1]) return DGX_ERR_ARG;
		const bool start = (e == 0 || p != pe[e - 1] || k != kind(e - 1));
		if (k == 0) {
			if (start) { order.push_back(p); e_off.push_back(e); }
			grp[e] = (int)order.size() - 1;
		} else {
			any_tri = true;
			if (start) { tri_p.push_back(p); tri_e0.push_back(e); }
			grp[e] = (int)tri_p.size() - 1;
		}
	}
	const int G = (int)order.size(), T = (int)tri_p.size();
	e_off.push_back(G ? (any_tri ? tri_e0[0] : E) : 0);
	tri_e0.push_back(E);
	auto npe = [&](int e) { return kind(e) ? (pe[e] + 1) * (pe[e] + 2) / 2 : (pe[e] + 1) * (pe[e] + 1); };
	std::vector<long long> uoff(E + 1, 0), xoff(E + 1, 0);
	for (int e = 0; e < E; ++e) {
		uoff[e + 1] = uoff[e] + (long long)npe(e) * nv;
		xoff[e + 1] = xoff[e] + (long long)npe(e) * 2;
	}
	std::vector<long long> z_off(G + 1, 0), tz_off(T + 1, 0);
	for (int g = 0; g < G; ++g)
		z_off[g + 1] = z_off[g] + (long long)(e_off[g + 1] - e_off[g]) * 4 * zblk * (order[g] + 1);
	tz_off[0] = z_off[G];
	for (int t = 0; t < T; ++t)
		tz_off[t + 1] = tz_off[t] + (long long)(tri_e0[t + 1] - tri_e0[t]) * 3 * zblk * (tri_p[t] + 1);
	// lifted-flux block of local face lf of element e (its own GL face points, [v][a]; Navier-Stokes: then y)
	auto zblock = [&](int e, int lf) -> long long {
		const int N = pe[e] + 1;
		if (kind(e)) return tz_off[grp[e]] + ((long long)(e - tri_e0[grp[e]]) * 3 + lf) * zblk * N;
		return z_off[grp[e]] + ((long long)(e - e_off[grp[e]]) * 4 + lf) * zblk * N;
	};
	// outward n J_f of element e's face lf at face coordinates s (its own map)
	auto side_nj = [&](int e, int lf, const double* s, int nq, double* out) {
		if (kind(e)) dgx::tri_face_nj(pe[e], lf, xyz + xoff[e], s, nq, out);
		else outward_nj(xyz + xoff[e], pe[e] + 1, lf, s, nq, out);
	};
	// split the face table: faces inside one group, faces between QUAD groups, faces with a TRI side
	std::vector<std::vector<int>> gfaces(G);
	std::vector<dgx::HpFace> hf;
	std::vector<dgx::MixFace> mf;
	std::vector<int> mface_elem;   // per mixed face: its (TRI) element when it is a rotating wall, else -1
	std::vector<double> nj, njh, mnj;   // n J_f of the face (lower-order side) and of the higher-order side's own face
	double xg[dgx::kHpMaxN], wg[dgx::kHpMaxN], tmp[2 * dgx::kHpMaxN];
	for (int f = 0; f < F; ++f) {
		const int* r = faces + 6 * f;
		const int eL = r[0], fL = r[1], eR = r[2], fR = r[3], ord = r[4], bc = r[5];
		if (eL < 0 || eL >= E || fL < 0 || fL >= (kind(eL) ? 3 : 4)) return DGX_ERR_ARG;
		// ord stays the reference's code in the group face tables (upload_uniform translates it); the faces
		// between orders use its permutation bits
		if ((bc == DGX_BC_INTERIOR && dgx::ord_bits(2, ord) < 0) || (bc != DGX_BC_INTERIOR && ord != -1 && ord != 0))
			return DGX_ERR_ARG;
		const int obits = (bc == DGX_BC_INTERIOR) ? dgx::ord_bits(2, ord) : 0;
		const int gL = grp[eL];
		auto push = [&](int g, int a0, int a1, int a2, int a3, int a4, int a5) {
			const int row[6] = { a0, a1, a2, a3, a4, a5 };
			gfaces[g].insert(gfaces[g].end(), row, row + 6);
		};
		if (bc == DGX_BC_INTERIOR && (eR < 0 || eR >= E || fR < 0 || fR >= (kind(eR) ? 3 : 4))) return DGX_ERR_ARG;
		// boundary codes: Euler takes Riemann and slip walls, Navier-Stokes Riemann and the no-slip walls
		if (bc != DGX_BC_INTERIOR) {
			const bool noslip = bc == DGX_BC_NOSLIP_ADIABATIC || bc == DGX_BC_NOSLIP_DIABATIC || bc == DGX_BC_NOSLIP_ROTATING;
			if (bc != DGX_BC_RIEMANN && bc != DGX_BC_SLIPWALL && !noslip) return DGX_ERR_ARG;
			if ((c.cfg.pde == DGX_PDE_NAVIER_STOKES) ? bc == DGX_BC_SLIPWALL : noslip) return DGX_ERR_UNSUPPORTED;
		}
		if (kind(eL) || (bc == DGX_BC_INTERIOR && kind(eR))) {
			// a TRI side: evaluated by the mixed face kernel, QUAD sides read its lifted flux
			dgx::MixFace x{};
			const bool inter = (bc == DGX_BC_INTERIOR);
			x.uL = uoff[eL];
			x.zL = zblock(eL, fL);
			x.kL = kind(eL);
			x.NL = pe[eL] + 1;
			x.lfL = fL;
			x.eL = kind(eL) ? eL - tri_e0[gL] : eL - e_off[gL];
			x.tL = kind(eL) ? gL : -1;
			x.bc = bc;
			x.rev = (obits & 2) ? 1 : 0;
			if (!kind(eL)) push(gL, eL - e_off[gL], fL, -1, -1, 0, dgx::kBcNonconforming);
			if (inter) {
				x.uR = uoff[eR];
				x.zR = zblock(eR, fR);
				x.kR = kind(eR);
				x.NR = pe[eR] + 1;
				x.lfR = fR;
				x.eR = kind(eR) ? eR - tri_e0[grp[eR]] : eR - e_off[grp[eR]];
				x.tR = kind(eR) ? grp[eR] : -1;
				if (!kind(eR)) push(grp[eR], eR - e_off[grp[eR]], fR, -1, -1, 0, dgx::kBcNonconforming);
			} else {
				x.NR = x.NL;
			}
			mf.push_back(x);
			mface_elem.push_back(bc == DGX_BC_NOSLIP_ROTATING ? eL : -1);
			const int Nf = std::max(x.NL, x.NR);
			dgx::gauss_legendre(Nf, xg, wg);
			const size_t base = mnj.size();
			mnj.resize(base + 2 * dgx::kHpMaxN, 0.0);
			if (!inter || x.NL <= x.NR) {
				side_nj(eL, fL, xg, Nf, mnj.data() + base);
			} else {
				double sr[dgx::kHpMaxN];
				for (int q = 0; q < Nf; ++q) sr[q] = x.rev ? -xg[q] : xg[q];
				side_nj(eR, fR, sr, Nf, tmp);
				for (int q = 0; q < 2 * Nf; ++q) mnj[base + q] = -tmp[q];
			}
			continue;
		}
		if (bc != DGX_BC_INTERIOR) {
			push(gL, eL - e_off[gL], fL, -1, -1, ord, bc);
			continue;
		}
		const int gR = grp[eR];
		if (gL == gR) {
			push(gL, eL - e_off[gL], fL, eR - e_off[gL], fR, ord, bc);
			continue;
		}
		push(gL, eL - e_off[gL], fL, -1, -1, 0, dgx::kBcNonconforming);
		push(gR, eR - e_off[gR], fR, -1, -1, 0, dgx::kBcNonconforming);
		const int NL = order[gL] + 1, NR = order[gR] + 1, Nf = std::max(NL, NR);
		dgx::HpFace x{};
		x.uL = uoff[eL];
		x.uR = uoff[eR];
		x.zL = z_off[gL] + ((long long)(eL - e_off[gL]) * 4 + fL) * zblk * NL;
		x.zR = z_off[gR] + ((long long)(eR - e_off[gR]) * 4 + fR) * zblk * NR;
		x.info = NL | (NR << 3) | (fL << 6) | (fR << 8) | (obits << 10);
		x.eL = eL - e_off[gL];
		x.eR = eR - e_off[gR];
		hf.push_back(x);
		// face geometry from the lower-order side, in the left ordering, left outward
		dgx::gauss_legendre(Nf, xg, wg);
		// (the higher-order side's own, used by the mixed form to cancel that element's constant part exactly)
		const size_t base = nj.size();
		nj.resize(base + 2 * dgx::kHpMaxN, 0.0);
		njh.resize(base + 2 * dgx::kHpMaxN, 0.0);
		double* lo = (NL < NR) ? nj.data() + base : njh.data() + base;   // the left side's own geometry
		double* ro = (NL < NR) ? njh.data() + base : nj.data() + base;   // the right side's, left orientation
		outward_nj(xyz + xoff[eL], NL, fL, xg, Nf, lo);
		outward_nj(xyz + xoff[eR], NR, fR, xg, Nf, tmp);
		for (int q = 0; q < Nf; ++q) {
			const int qr = (obits & 2) ? Nf - 1 - q : q;
			ro[2 * q] = -tmp[2 * qr];
			ro[2 * q + 1] = -tmp[2 * qr + 1];
		}
		// on a mesh with TRI elements, the same face for the FP64 face kernels, which then take every face
		// between orders or types by Nf (on QUAD-only meshes the per-pair kernels measured faster)
		if (!any_tri) continue;
		dgx::MixFace m{};
		m.uL = x.uL;
		m.uR = x.uR;
		m.zL = x.zL;
		m.zR = x.zR;
		m.NL = NL;
		m.NR = NR;
		m.lfL = fL;
		m.lfR = fR;
		m.rev = (obits & 2) ? 1 : 0;
		m.bc = DGX_BC_INTERIOR;
		m.eL = x.eL;
		m.eR = x.eR;
		m.tL = m.tR = -1;
		mf.push_back(m);
		mface_elem.push_back(-1);
		mnj.insert(mnj.end(), nj.begin() + base, nj.begin() + base + 2 * dgx::kHpMaxN);
	}
	// faces with a TRI side sorted by (Nf, kL, NL, kR, NR) (stable): one launch per Nf, waves of one shape
	std::vector<dgx::MixRun> mruns;
	{
		auto mkey = [&](const dgx::MixFace& x) {
			return ((std::max(x.NL, x.NR) * 2 + x.kL) * 8 + x.NL) * 16 + x.kR * 8 + x.NR;
		};
		std::vector<int> mp(mf.size());
		for (size_t i = 0; i < mf.size(); ++i) mp[i] = (int)i;
		std::stable_sort(mp.begin(), mp.end(), [&](int a, int b) { return mkey(mf[a]) < mkey(mf[b]); });
		std::vector<dgx::MixFace> ms(mf.size());
		std::vector<double> mn(mnj.size());
		std::vector<int> me(mf.size());
		for (size_t i = 0; i < mf.size(); ++i) {
			ms[i] = mf[mp[i]];
			me[i] = mface_elem[mp[i]];
			const size_t src = (size_t)mp[i] * 2 * dgx::kHpMaxN, dst = i * 2 * dgx::kHpMaxN;
			std::copy(mnj.begin() + src, mnj.begin() + src + 2 * dgx::kHpMaxN, mn.begin() + dst);
			const int Nf = std::max(ms[i].NL, ms[i].NR);
			if (mruns.empty() || mruns.back().Nf != Nf) mruns.push_back(dgx::MixRun{ Nf, (int)i, 0 });
			mruns.back().count += 1;
		}
		mf.swap(ms);
		mnj.swap(mn);
		mface_elem.swap(me);
	}
	// faces sorted by their (NL, NR) pair (stable): one kernel instance per pair, every wave on one pair
	std::vector<int> fperm(hf.size());
	for (size_t i = 0; i < hf.size(); ++i) fperm[i] = (int)i;
	auto pair_key = [&](int i) { return (hf[i].info & 7) * 8 + ((hf[i].info >> 3) & 7); };
	std::stable_sort(fperm.begin(), fperm.end(), [&](int a, int b) { return pair_key(a) < pair_key(b); });
	{
		std::vector<dgx::HpFace> hs(hf.size());
		std::vector<double> ns(nj.size()), nh(njh.size());
		for (size_t i = 0; i < hf.size(); ++i) {
			hs[i] = hf[fperm[i]];
			const size_t src = (size_t)fperm[i] * 2 * dgx::kHpMaxN, dst = i * 2 * dgx::kHpMaxN;
			std::copy(nj.begin() + src, nj.begin() + src + 2 * dgx::kHpMaxN, ns.begin() + dst);
			std::copy(njh.begin() + src, njh.begin() + src + 2 * dgx::kHpMaxN, nh.begin() + dst);
		}
		hf.swap(hs);
		nj.swap(ns);
		njh.swap(nh);
	}
	free_mesh(c);
	c.hp = new (std::nothrow) dgx::HpData();
	if (!c.hp) return DGX_ERR_ARG;
	dgx::HpData& H = *c.hp;
	for (size_t i = 0; i < hf.size(); ++i) {
		const int NL = hf[i].info & 7, NR = (hf[i].info >> 3) & 7;
		if (H.pairs.empty() || H.pairs.back().NL != NL || H.pairs.back().NR != NR)
			H.pairs.push_back(dgx::HpPair{ NL, NR, (int)i, 0 });
		H.pairs.back().count += 1;
	}
	H.order = order;
	H.e_off = e_off;
	H.E = E;
	H.F = F;
	H.Fnc = (int)hf.size();
	H.ndof = uoff[E];
	H.nz = tz_off[T];
	for (int g = 0; g < G; ++g) {
		H.u_off.push_back(uoff[e_off[g]]);
		H.z_off.push_back(z_off[g]);
	}
	DGX_TRY(hipMalloc(&H.z64, sizeof(double) * (size_t)std::max(H.nz, 1LL)));
	DGX_TRY(hipMalloc(&H.z32, sizeof(float) * (size_t)std::max(H.nz, 1LL)));
	DGX_TRY(hipMalloc(&H.faces, sizeof(dgx::HpFace) * std::max<size_t>(hf.size(), 1)));
	DGX_TRY(hipMalloc(&H.nJ, sizeof(double) * std::max<size_t>(nj.size(), 1)));
	DGX_TRY(hipMalloc(&H.nJ32, sizeof(float) * std::max<size_t>(nj.size(), 1)));
	DGX_TRY(hipMalloc(&H.nJhi, sizeof(double) * std::max<size_t>(njh.size(), 1)));
	DGX_TRY(hipMalloc(&H.ops64, sizeof(dgx::HpOps<double>)));
	DGX_TRY(hipMalloc(&H.ops32, sizeof(dgx::HpOps<float>)));
	if (!hf.empty()) {
		std::vector<float> nj32(nj.begin(), nj.end());
		DGX_TRY(hipMemcpy(H.faces, hf.data(), sizeof(dgx::HpFace) * hf.size(), hipMemcpyHostToDevice));
		DGX_TRY(hipMemcpy(H.nJ, nj.data(), sizeof(double) * nj.size(), hipMemcpyHostToDevice));
		DGX_TRY(hipMemcpy(H.nJ32, nj32.data(), sizeof(float) * nj32.size(), hipMemcpyHostToDevice));
		DGX_TRY(hipMemcpy(H.nJhi, njh.data(), sizeof(double) * njh.size(), hipMemcpyHostToDevice));
	}
	{
		dgx::HpOps<double>* o64 = new dgx::HpOps<double>;
		dgx::HpOps<float>* o32 = new dgx::HpOps<float>;
		build_hp_ops(*o64);
		const double* s64 = &o64->Ig[0][0][0];
		float* d32 = &o32->Ig[0][0][0];
		for (size_t i = 0; i < sizeof(dgx::HpOps<double>) / sizeof(double); ++i) d32[i] = (float)s64[i];
		const hipError_t e1 = hipMemcpy(H.ops64, o64, sizeof(*o64), hipMemcpyHostToDevice);
		const hipError_t e2 = hipMemcpy(H.ops32, o32, sizeof(*o32), hipMemcpyHostToDevice);
		delete o64;
		delete o32;
		if (e1 != hipSuccess || e2 != hipSuccess) return DGX_ERR_HIP;
	}
	// TRI groups: operators of every order, each element's metrics and det J at the cubature; the mixed faces
	if (T > 0) {
		std::vector<dgx::TriOps<double>> tops(dgx::kTriMaxP + 1);
		for (int p = 1; p <= dgx::kTriMaxP; ++p)
			if (!dgx::build_tri_ops(p, tops[p])) return DGX_ERR_UNSUPPORTED;
		DGX_TRY(hipMalloc(&H.tri_ops64, sizeof(dgx::TriOps<double>) * tops.size()));
		DGX_TRY(hipMemcpy(H.tri_ops64, tops.data(), sizeof(dgx::TriOps<double>) * tops.size(), hipMemcpyHostToDevice));
		{
			// the FP32 operators: every table rounded (the tables precede the sizes in TriOps)
			std::vector<dgx::TriOps<float>> t32(tops.size());
			constexpr size_t nt = offsetof(dgx::TriOps<double>, p) / sizeof(double);
			static_assert(offsetof(dgx::TriOps<float>, p) == nt * sizeof(float), "TriOps tables are contiguous");
			for (size_t k = 0; k < tops.size(); ++k) {
				const double* src = &tops[k].I[0];
				float* dst = &t32[k].I[0];
				for (size_t i = 0; i < nt; ++i) dst[i] = (float)src[i];
				t32[k].p = tops[k].p;
				t32[k].Np = tops[k].Np;
				t32[k].Nc = tops[k].Nc;
				t32[k].N = tops[k].N;
			}
			DGX_TRY(hipMalloc(&H.tri_ops32, sizeof(dgx::TriOps<float>) * t32.size()));
			DGX_TRY(hipMemcpy(H.tri_ops32, t32.data(), sizeof(dgx::TriOps<float>) * t32.size(), hipMemcpyHostToDevice));
		}
		for (int t = 0; t < T; ++t) {
			dgx::TriGroup tg;
			tg.p = tri_p[t];
			tg.e0 = tri_e0[t];
			tg.E = tri_e0[t + 1] - tri_e0[t];
			tg.u_off = uoff[tg.e0];
			tg.z_off = tz_off[t];
			const dgx::TriOps<double>& o = tops[tg.p];
			std::vector<double> met((size_t)tg.E * 4 * o.Nc), dj((size_t)tg.E * o.Nc);
			for (int k = 0; k < tg.E; ++k) {
				dgx::tri_geometry(o, xyz + xoff[tg.e0 + k], &met[(size_t)k * 4 * o.Nc], &dj[(size_t)k * o.Nc]);
				for (int q = 0; q < o.Nc; ++q)
					if (!(dj[(size_t)k * o.Nc + q] > 0.0)) return DGX_ERR_ARG;   // clockwise or degenerate element
			}
			H.tri.push_back(tg);
			dgx::TriGroup& hg = H.tri.back();
			const int Np = o.Np, N = o.N;
			DGX_TRY(hipMalloc(&hg.met, sizeof(double) * met.size()));
			DGX_TRY(hipMalloc(&hg.detJ, sizeof(double) * dj.size()));
			DGX_TRY(hipMalloc(&hg.tr, sizeof(double) * (size_t)tg.E * 3 * nv * N));
			DGX_TRY(hipMemcpy(hg.met, met.data(), sizeof(double) * met.size(), hipMemcpyHostToDevice));
			DGX_TRY(upload_f32(&hg.met32, met.data(), met.size()));
			DGX_TRY(hipMemcpy(hg.detJ, dj.data(), sizeof(double) * dj.size(), hipMemcpyHostToDevice));
			{
				// M^-1 (the explicit path; BR2: also the face-trace operators K_f = Ifc_f M^-1 tw0_f of the
				// lifts).  An affine element (det J constant at the cubature) has the reference ones divided by
				// det J; if every element of the group is affine only 1 / det J is stored, else each element's
				// own matrices.  The BR2 volume gradient and its face traces are written per evaluation.
				// affine: the least-squares affine map of the reference nodes onto the element's nodes reproduces
				// every node to 1e-12 of the element's size (det J computed at the cubature from the nodes varies
				// by rounding alone up to ~6e-13 at p = 5, so it cannot decide); its det J is then exact
				bool affine = true;
				std::vector<double> jinv(tg.E), rst(2 * (size_t)Np);
				if (dgx_reference_nodes(1, tg.p, rst.data(), nullptr, nullptr) != DGX_OK) return DGX_ERR_ARG;
				double nm[3][3] = {};   // normal equations of the fit in (r, s, 1)
				for (int k = 0; k < Np; ++k) {
					const double b3[3] = { rst[2 * k], rst[2 * k + 1], 1.0 };
					for (int i = 0; i < 3; ++i)
						for (int j = 0; j < 3; ++j) nm[i][j] += b3[i] * b3[j];
				}
				const double det3 = nm[0][0] * (nm[1][1] * nm[2][2] - nm[1][2] * nm[2][1]) -
				                    nm[0][1] * (nm[1][0] * nm[2][2] - nm[1][2] * nm[2][0]) +
				                    nm[0][2] * (nm[1][0] * nm[2][1] - nm[1][1] * nm[2][0]);
				for (int k = 0; k < tg.E && affine; ++k) {
					const double* x = xyz + xoff[tg.e0 + k];   // 2 x Np
					double coef[2][3];
					for (int dd = 0; dd < 2; ++dd) {
						double rhs[3] = {};
						for (int m = 0; m < Np; ++m) {
							rhs[0] += rst[2 * m] * x[dd * Np + m];
							rhs[1] += rst[2 * m + 1] * x[dd * Np + m];
							rhs[2] += x[dd * Np + m];
						}
						for (int c3 = 0; c3 < 3; ++c3) {   // Cramer's rule on the 3 x 3 normal equations
							double M3[3][3];
							for (int i = 0; i < 3; ++i)
								for (int j = 0; j < 3; ++j) M3[i][j] = (j == c3) ? rhs[i] : nm[i][j];
							coef[dd][c3] = (M3[0][0] * (M3[1][1] * M3[2][2] - M3[1][2] * M3[2][1]) -
							                M3[0][1] * (M3[1][0] * M3[2][2] - M3[1][2] * M3[2][0]) +
							                M3[0][2] * (M3[1][0] * M3[2][1] - M3[1][1] * M3[2][0])) / det3;
						}
					}
					double size = 0.0, worst = 0.0;
					for (int m = 0; m < Np; ++m)
						for (int dd = 0; dd < 2; ++dd) {
							const double fit = coef[dd][0] * rst[2 * m] + coef[dd][1] * rst[2 * m + 1] + coef[dd][2];
							worst = std::max(worst, std::fabs(fit - x[dd * Np + m]));
							size = std::max(size, std::fabs(x[dd * Np + m] - x[dd * Np]));
						}
					const double dA = coef[0][0] * coef[1][1] - coef[0][1] * coef[1][0];
					affine = worst <= 1e-12 * size && dA > 0.0;
					jinv[k] = 1.0 / dA;
				}
				hg.affine = affine ? 1 : 0;
				if (affine) {
					DGX_TRY(hipMalloc(&hg.jinv, sizeof(double) * jinv.size()));
					DGX_TRY(hipMemcpy(hg.jinv, jinv.data(), sizeof(double) * jinv.size(), hipMemcpyHostToDevice));
					DGX_TRY(upload_f32(&hg.jinv32, jinv.data(), jinv.size()));
				} else {
					std::vector<double> mi((size_t)tg.E * Np * Np), K((size_t)tg.E * 3 * N * N);
					for (int k = 0; k < tg.E; ++k)
						if (!dgx::tri_mass_data(o, &dj[(size_t)k * o.Nc], &mi[(size_t)k * Np * Np], &K[(size_t)k * 3 * N * N]))
							return DGX_ERR_ARG;
					DGX_TRY(hipMalloc(&hg.minv, sizeof(double) * mi.size()));
					DGX_TRY(hipMalloc(&hg.K, sizeof(double) * K.size()));
					DGX_TRY(hipMemcpy(hg.minv, mi.data(), sizeof(double) * mi.size(), hipMemcpyHostToDevice));
					DGX_TRY(hipMemcpy(hg.K, K.data(), sizeof(double) * K.size(), hipMemcpyHostToDevice));
					DGX_TRY(upload_f32(&hg.minv32, mi.data(), mi.size()));
					DGX_TRY(upload_f32(&hg.K32, K.data(), K.size()));
				}
				if (c.cfg.pde == DGX_PDE_NAVIER_STOKES) {
					DGX_TRY(hipMalloc(&hg.gv, sizeof(double) * (size_t)tg.E * 2 * nv * Np));
					DGX_TRY(hipMalloc(&hg.gvt, sizeof(double) * (size_t)tg.E * 3 * 2 * nv * N));
				}
			}
		}
	}
	// rotating walls on TRI faces: the coordinates of the face's GL points (its own, the face having one side)
	if (std::any_of(mface_elem.begin(), mface_elem.end(), [](int e) { return e >= 0; })) {
		std::vector<double> xy(mf.size() * dgx::kHpMaxN * 2, 0.0);
		std::vector<dgx::TriOps<double>> tcache(dgx::kTriMaxP + 1);
		std::vector<bool> built(dgx::kTriMaxP + 1, false);
		for (size_t i = 0; i < mf.size(); ++i) {
			const int e = mface_elem[i];
			if (e < 0) continue;
			if (!built[pe[e]]) {
				if (!dgx::build_tri_ops(pe[e], tcache[pe[e]])) return DGX_ERR_UNSUPPORTED;
				built[pe[e]] = true;
			}
			const dgx::TriOps<double>& to = tcache[pe[e]];
			const double* xe = xyz + xoff[e];
			const int N = pe[e] + 1, Np = to.Np;
			for (int q = 0; q < N; ++q) {
				double px = 0.0, py = 0.0;
				for (int m = 0; m < Np; ++m) {
					const double l = to.Ifc[mf[i].lfL][q * Np + m];
					px += l * xe[m];
					py += l * xe[Np + m];
				}
				xy[(i * dgx::kHpMaxN + q) * 2] = px;
				xy[(i * dgx::kHpMaxN + q) * 2 + 1] = py;
			}
		}
		DGX_TRY(hipMalloc(&H.mxy, sizeof(double) * xy.size()));
		DGX_TRY(hipMemcpy(H.mxy, xy.data(), sizeof(double) * xy.size(), hipMemcpyHostToDevice));
	}
	H.Fmix = (int)mf.size();
	H.mruns = mruns;
	if (!mf.empty()) {
		DGX_TRY(hipMalloc(&H.mfaces, sizeof(dgx::MixFace) * mf.size()));
		DGX_TRY(hipMalloc(&H.mnJ, sizeof(double) * mnj.size()));
		DGX_TRY(hipMemcpy(H.mfaces, mf.data(), sizeof(dgx::MixFace) * mf.size(), hipMemcpyHostToDevice));
		DGX_TRY(hipMemcpy(H.mnJ, mnj.data(), sizeof(double) * mnj.size(), hipMemcpyHostToDevice));
		DGX_TRY(upload_f32(&H.mnJ32, mnj.data(), mnj.size()));
	}
	// one uniform-order context per QUAD group
	for (int g = 0; g < G; ++g) {
		dgx_config cfg = c.cfg;
		cfg.p = order[g];
		dgx_ctx* sub = nullptr;
		dgx_status st = dgx_create(&sub, &cfg);
		if (st != DGX_OK) return st;
		H.groups.push_back(sub);
		sub->c.stream = c.stream;
		dgx_mesh sm{};
		sm.E = e_off[g + 1] - e_off[g];
		sm.F = (int)(gfaces[g].size() / 6);
		sm.xyz = xyz + xoff[e_off[g]];
		sm.faces = gfaces[g].empty() ? nullptr : gfaces[g].data();
		st = upload_uniform(sub, &sm, true);
		if (st != DGX_OK) return st;
		sub->c.hp_z64 = H.z64 + H.z_off[g];
		sub->c.hp_z32 = H.z32 + H.z_off[g];
	}
	// side streams, and the groups and face pairs spread over them longest first (cost ~ items x N^2)
	for (int k = 0; k < dgx::HpData::kLanes; ++k) {
		DGX_TRY(hipStreamCreateWithFlags(&H.lane[k], hipStreamNonBlocking));
		DGX_TRY(hipEventCreateWithFlags(&H.join_ev[k], hipEventDisableTiming));
	}
	DGX_TRY(hipEventCreateWithFlags(&H.fork_ev, hipEventDisableTiming));
	auto balance = [](const std::vector<double>& cost) {
		std::vector<int> idx(cost.size()), lane(cost.size(), 0);
		for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
		std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return cost[a] > cost[b]; });
		double load[dgx::HpData::kLanes] = {};
		for (int i : idx) {
			int k = 0;
			for (int j = 1; j < dgx::HpData::kLanes; ++j)
				if (load[j] < load[k]) k = j;
			lane[i] = k;
			load[k] += cost[i];
		}
		return lane;
	};
	// the QUAD and TRI groups of one phase share the streams: one balance over both (a TRI element's dense
	// products cost about Np Nc / 3 against N^2 per QUAD element, from the measured kernel times)
	std::vector<double> gcost(G + T), pcost(H.pairs.size());
	for (int g = 0; g < G; ++g) gcost[g] = (double)(e_off[g + 1] - e_off[g]) * (order[g] + 1) * (order[g] + 1);
	for (int t = 0; t < T; ++t) {
		const int p = tri_p[t], np = (p + 1) * (p + 2) / 2, nc = p == 1 ? 3 : p == 2 ? 6 : p == 3 ? 12 : p == 4 ? 16 : 25;
		gcost[G + t] = (double)(tri_e0[t + 1] - tri_e0[t]) * np * nc / 3.0;
	}
	for (size_t i = 0; i < H.pairs.size(); ++i) {
		const int n = std::max(H.pairs[i].NL, H.pairs[i].NR);
		pcost[i] = (double)H.pairs[i].count * n * n;
	}
	const std::vector<int> lanes = balance(gcost);
	H.group_lane.assign(lanes.begin(), lanes.begin() + G);
	H.tri_lane.assign(lanes.begin() + G, lanes.end());
	H.pair_lane = balance(pcost);
	std::vector<double> mcost(H.mruns.size());
	for (size_t i = 0; i < H.mruns.size(); ++i) mcost[i] = (double)H.mruns[i].count * H.mruns[i].Nf * H.mruns[i].Nf;
	H.mrun_lane = balance(mcost);
	H.ready = true;
	return DGX_OK;
}

// Sutherland's viscosity is built for the uniform-order QUAD kernels only
static bool sutherland_ns(const dgx_ctx* h)
{
	return h && h->c.cfg.pde == DGX_PDE_NAVIER_STOKES && h->c.cfg.viscosity_type == DGX_VISC_SUTHERLAND;
}

dgx_status dgx_upload_mesh_hp(dgx_ctx* h, const dgx_mesh_hp* m)
{
	if (sutherland_ns(h)) return DGX_ERR_UNSUPPORTED;
	const dgx_status st = m ? upload_hp_impl(h, m->E, m->F, nullptr, m->p, m->xyz, m->faces) : DGX_ERR_ARG;
	// a failure after the previous mesh was released leaves no half-built variable-order state behind
	if (st != DGX_OK && h && h->c.hp && !h->c.hp->ready) free_mesh(h->c);
	return st;
}

dgx_status dgx_upload_mesh_mixed(dgx_ctx* h, const dgx_mesh_mixed* m)
{
	if (sutherland_ns(h)) return DGX_ERR_UNSUPPORTED;
	const dgx_status st = (m && m->type) ? upload_hp_impl(h, m->E, m->F, m->type, m->p, m->xyz, m->faces)
	                                     : DGX_ERR_ARG;
	if (st != DGX_OK && h && h->c.hp && !h->c.hp->ready) free_mesh(h->c);
	return st;
}

dgx_status dgx_get_geometry(dgx_ctx* h, double* met, double* dj, double* nJ)
{
	if (!h) return DGX_ERR_ARG;
	Ctx& c = h->c;
	if (!c.mesh_ready) return c.hp ? DGX_ERR_UNSUPPORTED : DGX_ERR_NO_MESH;
	DGX_TRY(hipStreamSynchronize(c.stream));
	if (met) DGX_TRY(hipMemcpy(met, c.metrics_vc, sizeof(double) * (size_t)c.E * c.d * c.d * c.Nc, hipMemcpyDeviceToHost));
	if (dj) DGX_TRY(hipMemcpy(dj, c.detJ_vc, sizeof(double) * (size_t)c.E * c.Nc, hipMemcpyDeviceToHost));
	if (nJ && c.F > 0) DGX_TRY(hipMemcpy(nJ, c.nJ_fc, sizeof(double) * (size_t)c.F * c.Nfc * c.d, hipMemcpyDeviceToHost));
	return DGX_OK;
}

dgx_status dgx_compute_rhs(dgx_ctx* h, const double* u, double* r)
{
	if (!h || !u || !r) return DGX_ERR_ARG;
	if (h->c.hp) return hp_rhs64(h->c, u, r, dgx::MODE_RHS, nullptr);
	if (!h->c.mesh_ready) return h->c.hp ? DGX_ERR_UNSUPPORTED : DGX_ERR_NO_MESH;
	return hip_status(dgx::launch_rhs64(h->c, u, r, dgx::MODE_RHS, nullptr));
}

dgx_status dgx_compute_rhs_f32(dgx_ctx* h, const float* u, float* r)
{
	if (!h || !u || !r) return DGX_ERR_ARG;
	if (h->c.hp) return hp_rhs32(h->c, u, r);
	if (!h->c.mesh_ready) return h->c.hp ? DGX_ERR_UNSUPPORTED : DGX_ERR_NO_MESH;
	if (h->c.cfg.pde == DGX_PDE_NAVIER_STOKES) return hip_status(dgx::launch_rhs_ns32(h->c, u, r));
	return hip_status(dgx::launch_rhs32(h->c, u, r));
}

dgx_status dgx_compute_rhs_mixed(dgx_ctx* h, const double* u, float* r)
{
	if (!h || !u || !r) return DGX_ERR_ARG;
	if (h->c.hp) return hp_rhs_mixed(h->c, u, r);
	if (!h->c.mesh_ready) return h->c.hp ? DGX_ERR_UNSUPPORTED : DGX_ERR_NO_MESH;
	if (h->c.cfg.pde == DGX_PDE_NAVIER_STOKES) return hip_status(dgx::launch_rhs_ns_mixed(h->c, u, r));
	if (h->c.cfg.pde != DGX_PDE_EULER) return DGX_ERR_UNSUPPORTED;
	return hip_status(dgx::launch_rhs_mixed(h->c, u, r));
}

static dgx_status read_max(Ctx& c, double* out)
{
	unsigned long long bits = 0;
	DGX_TRY(hipMemcpyAsync(&bits, c.dmax, sizeof(bits), hipMemcpyDeviceToHost, c.stream));
	DGX_TRY(hipStreamSynchronize(c.stream));
	double v;
	std::memcpy(&v, &bits, sizeof(v));
	*out = v;
	return (v != v || v > 1.7e308) ? DGX_ERR_NONFINITE : DGX_OK;
}

static dgx_status rhs_part(dgx_ctx* h, const double* u, double* r, bool volume)
{
	if (!h || !u || !r) return DGX_ERR_ARG;
	Ctx& c = h->c;
	if (c.hp) return DGX_ERR_UNSUPPORTED;
	if (!c.mesh_ready) return DGX_ERR_NO_MESH;
	if (c.cfg.pde == DGX_PDE_NAVIER_STOKES) return DGX_ERR_UNSUPPORTED;
	return hip_status(dgx::launch_rhs_part64(c, u, r, volume));
}

dgx_status dgx_compute_volume_rhs(dgx_ctx* h, const double* u, double* r) { return rhs_part(h, u, r, true); }

dgx_status dgx_compute_face_rhs(dgx_ctx* h, const double* u, double* r) { return rhs_part(h, u, r, false); }

dgx_status dgx_compute_rhs_explicit(dgx_ctx* h, const double* u, double* r, double* max_abs_r)
{
	if (!h || !u || !r) return DGX_ERR_ARG;
	Ctx& c = h->c;
	if (c.hp) {
		DGX_TRY(dgx::launch_fill_u64(c.stream, c.dmax, 1, 0ull));
		const dgx_status st = hp_rhs64(c, u, r, dgx::MODE_RHS_MINV, c.dmax);
		if (st != DGX_OK) return st;
		return max_abs_r ? read_max(c, max_abs_r) : DGX_OK;
	}
	if (!c.mesh_ready) return DGX_ERR_NO_MESH;
	DGX_TRY(dgx::launch_fill_u64(c.stream, c.dmax, 1, 0ull));
	DGX_TRY(dgx::launch_rhs64(c, u, r, dgx::MODE_RHS_MINV, c.dmax));
	if (max_abs_r) return read_max(c, max_abs_r);
	return DGX_OK;
}

dgx_status dgx_apply_minv(dgx_ctx* h, double* r)
{
	if (!h || !r) return DGX_ERR_ARG;
	if (!h->c.mesh_ready) return h->c.hp ? DGX_ERR_UNSUPPORTED : DGX_ERR_NO_MESH;
	return hip_status(dgx::launch_minv(h->c, r));
}

dgx_status dgx_enforce_positivity(dgx_ctx* h, double* u)
{
	if (!h || !u) return DGX_ERR_ARG;
	Ctx& c = h->c;
	if (!c.mesh_ready) return c.hp ? DGX_ERR_UNSUPPORTED : DGX_ERR_NO_MESH;
	if (c.cfg.pde == DGX_PDE_ADVECTION) return DGX_OK;  // test_case.c:63-77: Euler/NS only
	DGX_TRY(dgx::launch_fill_i32(c.stream, c.dstatus, 1, 0));
	DGX_TRY(dgx::launch_positivity(c, u, c.dstatus));
	int st = 0;
	DGX_TRY(hipMemcpyAsync(&st, c.dstatus, sizeof(int), hipMemcpyDeviceToHost, c.stream));
	DGX_TRY(hipStreamSynchronize(c.stream));
	return st ? DGX_ERR_POSITIVITY : DGX_OK;
}

dgx_status dgx_ssp_rk3_step(dgx_ctx* h, double* u, double dt, int positivity, double* max_abs_r)
{
	if (!h || !u) return DGX_ERR_ARG;
	Ctx& c = h->c;
	if (!c.mesh_ready && !c.hp) return DGX_ERR_NO_MESH;
	const size_t n = ndof(c);
	if (!c.u0) DGX_TRY(hipMalloc(&c.u0, sizeof(double) * n));
	if (!c.rtmp) DGX_TRY(hipMalloc(&c.rtmp, sizeof(double) * n));
	const bool pos = positivity && c.cfg.pde != DGX_PDE_ADVECTION;
	if (pos && c.hp && !c.hp->tri.empty()) return DGX_ERR_UNSUPPORTED;   // the limiter is built for QUAD/HEX only
	if (pos) DGX_TRY(dgx::launch_fill_i32(c.stream, c.dstatus, 1, 0));
	for (int rk = 0; rk < 3; ++rk) {
		DGX_TRY(dgx::launch_fill_u64(c.stream, c.dmax, 1, 0ull));
		if (c.hp) {
			// every group's residual from the stage's state, then every group's stage update
			const dgx_status st = hp_rhs64(c, u, c.rtmp, dgx::MODE_RHS_MINV, c.dmax);
			if (st != DGX_OK) return st;
			dgx::HpData& H = *c.hp;
			for (size_t g = 0; g < H.groups.size(); ++g) {
				Ctx& s = H.groups[g]->c;
				if (!s.u0) DGX_TRY(hipMalloc(&s.u0, sizeof(double) * ndof(s)));
				DGX_TRY(dgx::launch_rk_stage(s, u + H.u_off[g], c.rtmp + H.u_off[g], dt, rk));
				if (pos) DGX_TRY(dgx::launch_positivity(s, u + H.u_off[g], c.dstatus));
			}
			for (const dgx::TriGroup& t : H.tri) {   // TRI groups: the stage copy in the mesh-wide buffer
				const size_t nt = (size_t)t.E * c.nv * ((t.p + 1) * (t.p + 2) / 2);
				DGX_TRY(dgx::launch_rk_stage_n(c.stream, u + t.u_off, c.u0 + t.u_off, c.rtmp + t.u_off, dt, rk, nt));
			}
			continue;
		}
		DGX_TRY(dgx::launch_rhs64(c, u, c.rtmp, dgx::MODE_RHS_MINV, c.dmax));
		DGX_TRY(dgx::launch_rk_stage(c, u, c.rtmp, dt, rk));
		if (pos) DGX_TRY(dgx::launch_positivity(c, u, c.dstatus));
	}
	dgx_status st = DGX_OK;
	if (max_abs_r) st = read_max(c, max_abs_r);
	if (pos) {
		int bad = 0;
		DGX_TRY(hipMemcpyAsync(&bad, c.dstatus, sizeof(int), hipMemcpyDeviceToHost, c.stream));
		DGX_TRY(hipStreamSynchronize(c.stream));
		if (bad) return DGX_ERR_POSITIVITY;
	}
	return st;
}

}  // extern "C"

struct dgx_graph {
	dgx_ctx* ctx = nullptr;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	bool rk3 = false, positivity = false;
};

namespace {

// Records what `record` launches on the context's stream into g (a private non-blocking stream stands in for
// the context's stream during the capture, so capture also works for a context on the null stream).
template <typename F>
dgx_status capture_graph(dgx_ctx* h, dgx_graph* g, F record)
{
	Ctx& c = h->c;
	hipStream_t cs = nullptr;
	DGX_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
	const hipStream_t saved = c.stream;
	c.stream = cs;
	hipGraph_t graph = nullptr;
	hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
	if (e == hipSuccess) {
		const hipError_t er = record();
		e = hipStreamEndCapture(cs, &graph);
		if (e == hipSuccess) e = er;
	}
	c.stream = saved;
	(void)hipStreamDestroy(cs);
	if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
	if (e != hipSuccess) {
		if (graph) (void)hipGraphDestroy(graph);
		return e == hipErrorInvalidValue ? DGX_ERR_UNSUPPORTED : DGX_ERR_HIP;
	}
	g->graph = graph;
	g->ctx = h;
	return DGX_OK;
}

}  // namespace

extern "C" {

dgx_status dgx_graph_create_rhs(dgx_ctx* h, int kind, const void* u, void* r, dgx_graph** out)
{
	if (!h || !u || !r || !out || kind < DGX_GRAPH_RHS_FP64 || kind > DGX_GRAPH_RHS_MIXED) return DGX_ERR_ARG;
	Ctx& c = h->c;
	if (c.hp) {
		// variable-order mesh: the face kernel between orders and every order group's launches in one graph
		dgx_graph* g = new (std::nothrow) dgx_graph();
		if (!g) return DGX_ERR_HIP;
		const dgx_status st = capture_graph(h, g, [&]() -> hipError_t {
			const dgx_status s = (kind == DGX_GRAPH_RHS_FP64)
			    ? hp_rhs64(c, static_cast<const double*>(u), static_cast<double*>(r), dgx::MODE_RHS, nullptr)
			    : (kind == DGX_GRAPH_RHS_FP32)
			    ? hp_rhs32(c, static_cast<const float*>(u), static_cast<float*>(r))
			    : hp_rhs_mixed(c, static_cast<const double*>(u), static_cast<float*>(r));
			return s == DGX_OK ? hipSuccess : hipErrorLaunchFailure;
		});
		if (st != DGX_OK) {
			delete g;
			return st;
		}
		*out = g;
		return DGX_OK;
	}
	if (!c.mesh_ready) return DGX_ERR_NO_MESH;
	if (kind == DGX_GRAPH_RHS_MIXED && c.cfg.pde == DGX_PDE_ADVECTION) return DGX_ERR_UNSUPPORTED;
	dgx_graph* g = new (std::nothrow) dgx_graph();
	if (!g) return DGX_ERR_HIP;
	const dgx_status st = capture_graph(h, g, [&]() -> hipError_t {
		if (kind == DGX_GRAPH_RHS_FP64)
			return dgx::launch_rhs64(c, static_cast<const double*>(u), static_cast<double*>(r), dgx::MODE_RHS, nullptr);
		if (kind == DGX_GRAPH_RHS_FP32 && c.cfg.pde == DGX_PDE_NAVIER_STOKES)
			return dgx::launch_rhs_ns32(c, static_cast<const float*>(u), static_cast<float*>(r));
		if (kind == DGX_GRAPH_RHS_FP32)
			return dgx::launch_rhs32(c, static_cast<const float*>(u), static_cast<float*>(r));
		if (c.cfg.pde == DGX_PDE_NAVIER_STOKES)
			return dgx::launch_rhs_ns_mixed(c, static_cast<const double*>(u), static_cast<float*>(r));
		return dgx::launch_rhs_mixed(c, static_cast<const double*>(u), static_cast<float*>(r));
	});
	if (st != DGX_OK) {
		delete g;
		return st;
	}
	*out = g;
	return DGX_OK;
}

dgx_status dgx_graph_create_rk3(dgx_ctx* h, double* u, double dt, int positivity, dgx_graph** out)
{
	if (!h || !u || !out) return DGX_ERR_ARG;
	Ctx& c = h->c;
	if (!c.mesh_ready) return c.hp ? DGX_ERR_UNSUPPORTED : DGX_ERR_NO_MESH;
	const size_t n = ndof(c);
	if (!c.u0) DGX_TRY(hipMalloc(&c.u0, sizeof(double) * n));     // stage buffers exist before the capture
	if (!c.rtmp) DGX_TRY(hipMalloc(&c.rtmp, sizeof(double) * n));
	dgx_graph* g = new (std::nothrow) dgx_graph();
	if (!g) return DGX_ERR_HIP;
	g->rk3 = true;
	g->positivity = positivity && c.cfg.pde != DGX_PDE_ADVECTION;
	const dgx_status st = capture_graph(h, g, [&]() -> hipError_t {
		for (int rk = 0; rk < 3; ++rk) {
			DGX_HIP_CHECK(dgx::launch_fill_u64(c.stream, c.dmax, 1, 0ull));
			DGX_HIP_CHECK(dgx::launch_rhs64(c, u, c.rtmp, dgx::MODE_RHS_MINV, c.dmax));
			DGX_HIP_CHECK(dgx::launch_rk_stage(c, u, c.rtmp, dt, rk));
			if (g->positivity) DGX_HIP_CHECK(dgx::launch_positivity(c, u, c.dstatus));
		}
		return hipSuccess;
	});
	if (st != DGX_OK) {
		delete g;
		return st;
	}
	// the limiter flag accumulates over replays until dgx_graph_sync reads and clears it
	if (g->positivity && dgx::launch_fill_i32(c.stream, c.dstatus, 1, 0) != hipSuccess) {
		(void)dgx_graph_destroy(g);
		return DGX_ERR_HIP;
	}
	*out = g;
	return DGX_OK;
}

dgx_status dgx_graph_launch(dgx_graph* g, int n)
{
	if (!g || !g->exec || n < 0) return DGX_ERR_ARG;
	for (int i = 0; i < n; ++i) DGX_TRY(hipGraphLaunch(g->exec, g->ctx->c.stream));
	return DGX_OK;
}

dgx_status dgx_graph_sync(dgx_graph* g, double* max_abs_r)
{
	if (!g || !g->exec) return DGX_ERR_ARG;
	Ctx& c = g->ctx->c;
	DGX_TRY(hipStreamSynchronize(c.stream));
	if (!g->rk3) return DGX_OK;
	dgx_status st = DGX_OK;
	if (max_abs_r) st = read_max(c, max_abs_r);
	if (g->positivity) {
		int bad = 0;
		DGX_TRY(hipMemcpyAsync(&bad, c.dstatus, sizeof(int), hipMemcpyDeviceToHost, c.stream));
		DGX_TRY(dgx::launch_fill_i32(c.stream, c.dstatus, 1, 0));
		DGX_TRY(hipStreamSynchronize(c.stream));
		if (bad) return DGX_ERR_POSITIVITY;
	}
	return st;
}

dgx_status dgx_graph_destroy(dgx_graph* g)
{
	if (!g) return DGX_OK;
	if (g->exec) (void)hipGraphExecDestroy(g->exec);
	if (g->graph) (void)hipGraphDestroy(g->graph);
	delete g;
	return DGX_OK;
}

dgx_status dgx_set_rhs_grid(dgx_ctx* h, int kind, int grid)
{
	if (!h || kind < DGX_GRAPH_RHS_FP64 || kind > DGX_TUNE_RHS_MINV || grid < 0) return DGX_ERR_ARG;
	h->c.rhs_grid[kind] = grid;
	return DGX_OK;
}

static hipError_t launch_rhs_kind(Ctx& c, int kind, const void* u, void* r)
{
	if (kind == dgx::TUNE_RHS64)
		return dgx::launch_rhs64(c, static_cast<const double*>(u), static_cast<double*>(r), dgx::MODE_RHS, nullptr);
	if (kind == dgx::TUNE_RHS32) return dgx::launch_rhs32(c, static_cast<const float*>(u), static_cast<float*>(r));
	if (kind == dgx::TUNE_MIXED) return dgx::launch_rhs_mixed(c, static_cast<const double*>(u), static_cast<float*>(r));
	return dgx::launch_rhs64(c, static_cast<const double*>(u), static_cast<double*>(r), dgx::MODE_RHS_MINV, nullptr);
}

dgx_status dgx_autotune_rhs(dgx_ctx* h, int kind, const void* u, void* r, int* grid_out, double* ms_out)
{
	if (!h || !u || !r || kind < DGX_GRAPH_RHS_FP64 || kind > DGX_TUNE_RHS_MINV) return DGX_ERR_ARG;
	Ctx& c = h->c;
	if (!c.mesh_ready) return c.hp ? DGX_ERR_UNSUPPORTED : DGX_ERR_NO_MESH;
	if (kind == DGX_GRAPH_RHS_MIXED && c.cfg.pde != DGX_PDE_EULER) return DGX_ERR_UNSUPPORTED;
	if (c.cfg.pde == DGX_PDE_NAVIER_STOKES) return DGX_ERR_UNSUPPORTED;   // its kernels have a fixed launch shape
	int dev = 0, ncu = 0;
	DGX_TRY(hipGetDevice(&dev));
	DGX_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
	// one workgroup per element block, then persistent grids, then the first again (the clocks may still be
	// ramping up during the first measurement)
	std::vector<int> cand = { 0 };
	for (int m : { 1, 2, 3, 4, 6, 8, 12, 16 }) cand.push_back(m * ncu);
	cand.push_back(0);
	hipEvent_t e0 = nullptr, e1 = nullptr;
	if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
		if (e0) (void)hipEventDestroy(e0);
		return DGX_ERR_HIP;
	}
	constexpr int kReps = 5;
	const int saved = c.rhs_grid[kind];
	int best = 0;
	float best_ms = 0.0f, full_ms = 0.0f;
	hipError_t err = hipSuccess;
	for (size_t k = 0; k < cand.size() && err == hipSuccess; ++k) {
		c.rhs_grid[kind] = cand[k];
		err = launch_rhs_kind(c, kind, u, r);   // warm-up
		if (err == hipSuccess) err = hipEventRecord(e0, c.stream);
		for (int rep = 0; rep < kReps && err == hipSuccess; ++rep) err = launch_rhs_kind(c, kind, u, r);
		if (err == hipSuccess) err = hipEventRecord(e1, c.stream);
		if (err == hipSuccess) err = hipEventSynchronize(e1);
		float ms = 0.0f;
		if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
		ms /= kReps;
		if (k == 0) full_ms = best_ms = ms;
		else if (cand[k] == 0) {
			full_ms = std::min(full_ms, ms);
			if (ms < best_ms) { best_ms = ms; best = 0; }
		} else if (ms < best_ms) { best_ms = ms; best = cand[k]; }
	}
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	if (err != hipSuccess) {
		c.rhs_grid[kind] = saved;
		return DGX_ERR_HIP;
	}
	// a persistent grid is kept only when it beats one-workgroup-per-block by more than 1%
	if (best != 0 && best_ms > 0.99f * full_ms) { best = 0; best_ms = full_ms; }
	c.rhs_grid[kind] = best;
	if (grid_out) *grid_out = best;
	if (ms_out) *ms_out = best_ms;
	return DGX_OK;
}

dgx_status dgx_max_abs(dgx_ctx* h, const double* x, long long n, double* out)
{
	if (!h || !x || !out || n < 0) return DGX_ERR_ARG;
	Ctx& c = h->c;
	DGX_TRY(dgx::launch_fill_u64(c.stream, c.dmax, 1, 0ull));
	DGX_TRY(dgx::launch_maxabs(c, x, (size_t)n, c.dmax));
	return read_max(c, out);
}

dgx_status dgx_convert_f64_f32(dgx_ctx* h, const double* src, float* dst, long long n)
{
	if (!h || !src || !dst || n < 0) return DGX_ERR_ARG;
	return hip_status(dgx::launch_convert(src, dst, (size_t)n, h->c.stream));
}

dgx_status dgx_convert_f32_f64(dgx_ctx* h, const float* src, double* dst, long long n)
{
	if (!h || !src || !dst || n < 0) return DGX_ERR_ARG;
	return hip_status(dgx::launch_convert_back(src, dst, (size_t)n, h->c.stream));
}

dgx_status dgx_assemble_diag_blocks(dgx_ctx* h, const double* u, const double* mass_scale, double* blocks)
{
	if (!h || !u || !blocks) return DGX_ERR_ARG;
	Ctx& c = h->c;
	if (!c.mesh_ready) return c.hp ? DGX_ERR_UNSUPPORTED : DGX_ERR_NO_MESH;
	if (c.d != 2 && c.d != 3) return DGX_ERR_UNSUPPORTED;   // 3-D: Euler and advection
	const int nb = c.Np * c.nv;
	for (int col = 0; col < nb; ++col) DGX_TRY(dgx::launch_diag_col(c, u, blocks, col));
	if (mass_scale) DGX_TRY(dgx::launch_sub_mass(c, blocks, mass_scale, 0.0));
	return DGX_OK;
}

dgx_status dgx_assemble_face_blocks(dgx_ctx* h, const double* u, double* blocks)
{
	if (!h || !u || !blocks) return DGX_ERR_ARG;
	Ctx& c = h->c;
	if (!c.mesh_ready) return c.hp ? DGX_ERR_UNSUPPORTED : DGX_ERR_NO_MESH;
	if (c.d != 2 && c.d != 3) return DGX_ERR_UNSUPPORTED;
	if (c.cfg.pde == DGX_PDE_NAVIER_STOKES) {
		if (c.d != 2) return DGX_ERR_UNSUPPORTED;
		// BR2: full blocks (the neighbour's whole state enters through its lifted gradient), one column per launch
		const int nb = c.Np * c.nv;
		for (int col = 0; col < nb; ++col) DGX_TRY(dgx::launch_ns_offd_col(c, u, blocks, col, false));
		return DGX_OK;
	}
	const int nfv = c.nv * (c.d == 2 ? c.N : c.N * c.N);   // face-node rows / columns of a face block
	double* cT = nullptr;
	DGX_TRY(hipMalloc(&cT, sizeof(double) * (size_t)c.E * 2 * c.d * nfv * nfv));
	hipError_t e = hipSuccess;
	for (int col = 0; col < nfv && e == hipSuccess; ++col) e = dgx::launch_offd_col(c, u, cT, col);
	if (e == hipSuccess) e = dgx::launch_expand_face_blocks(c, cT, blocks);
	const hipError_t es = hipStreamSynchronize(c.stream);
	(void)hipFree(cT);
	return hip_status(e != hipSuccess ? e : es);
}

namespace {

// the block pattern of the Newton matrix (constructor_nnz_dg_T, solve_dg_T.c:41-74): per element row, its own
// column and its face neighbours', ascending and without repeats; diag[e] and face[lf * E + e] give each
// element's and each local face's block slot (-1: a boundary face)
dgx_status bsr_pattern(Ctx& c, std::vector<int>& row_ptr, std::vector<int>& col, std::vector<int>& diag,
                       std::vector<int>& face)
{
	const int E = c.E, nf = 2 * c.d;
	std::vector<dgx::ElemFace> ef((size_t)E * nf);
	if (E > 0 && hipMemcpy(ef.data(), c.elem_faces, sizeof(dgx::ElemFace) * ef.size(), hipMemcpyDeviceToHost) != hipSuccess)
		return DGX_ERR_HIP;
	row_ptr.assign(E + 1, 0);
	col.clear();
	diag.assign(E, -1);
	face.assign((size_t)nf * E, -1);
	for (int e = 0; e < E; ++e) {
		std::vector<int> cs{ e };
		for (int lf = 0; lf < nf; ++lf)
			if (ef[(size_t)e * nf + lf].en >= 0) cs.push_back(ef[(size_t)e * nf + lf].en);
		std::sort(cs.begin(), cs.end());
		cs.erase(std::unique(cs.begin(), cs.end()), cs.end());
		const int base = (int)col.size();
		for (int k = 0; k < (int)cs.size(); ++k) {
			if (cs[k] == e) diag[e] = base + k;
			col.push_back(cs[k]);
		}
		for (int lf = 0; lf < nf; ++lf) {
			const int en = ef[(size_t)e * nf + lf].en;
			if (en >= 0) face[(size_t)lf * E + e] = base + (int)(std::lower_bound(cs.begin(), cs.end(), en) - cs.begin());
		}
		row_ptr[e + 1] = (int)col.size();
	}
	return DGX_OK;
}

}  // namespace

dgx_status dgx_jacobian_bsr_pattern(dgx_ctx* h, int* row_ptr, int* col_idx, long long* nnzb)
{
	if (!h || !nnzb) return DGX_ERR_ARG;
	Ctx& c = h->c;
	if (!c.mesh_ready) return c.hp ? DGX_ERR_UNSUPPORTED : DGX_ERR_NO_MESH;
	std::vector<int> rp, col, diag, face;
	const dgx_status ps = bsr_pattern(c, rp, col, diag, face);
	if (ps != DGX_OK) return ps;
	*nnzb = (long long)col.size();
	if (row_ptr) std::copy(rp.begin(), rp.end(), row_ptr);
	if (col_idx) std::copy(col.begin(), col.end(), col_idx);
	return DGX_OK;
}

dgx_status dgx_assemble_jacobian_bsr(dgx_ctx* h, const double* u, const double* mass_scale, double* values)
{
	if (!h || !u || !values) return DGX_ERR_ARG;
	Ctx& c = h->c;
	if (!c.mesh_ready) return c.hp ? DGX_ERR_UNSUPPORTED : DGX_ERR_NO_MESH;
	if (c.d != 2 && (c.d != 3 || c.cfg.pde == DGX_PDE_NAVIER_STOKES)) return DGX_ERR_UNSUPPORTED;
	std::vector<int> rp, col, diag, face;
	const dgx_status ps = bsr_pattern(c, rp, col, diag, face);
	if (ps != DGX_OK) return ps;
	const int E = c.E, nb = c.Np * c.nv, nb2 = nb * nb, nf = 2 * c.d;
	double *dblk = nullptr, *fblk = nullptr;
	int* slots = nullptr;
	dgx_status st = DGX_OK;
	do {
		if (hipMalloc(&dblk, sizeof(double) * (size_t)E * nb2) != hipSuccess ||
		    hipMalloc(&fblk, sizeof(double) * (size_t)E * nf * nb2) != hipSuccess ||
		    hipMalloc(&slots, sizeof(int) * (size_t)(nf + 1) * E) != hipSuccess) { st = DGX_ERR_HIP; break; }
		if (hipMemcpyAsync(slots, diag.data(), sizeof(int) * E, hipMemcpyHostToDevice, c.stream) != hipSuccess ||
		    hipMemcpyAsync(slots + E, face.data(), sizeof(int) * nf * (size_t)E, hipMemcpyHostToDevice, c.stream) != hipSuccess) {
			st = DGX_ERR_HIP;
			break;
		}
		if ((st = dgx_assemble_diag_blocks(h, u, mass_scale, dblk)) != DGX_OK) break;
		if ((st = dgx_assemble_face_blocks(h, u, fblk)) != DGX_OK) break;
		// the face blocks are added into their slots: every slot starts at zero
		for (size_t off = 0, n = col.size() * (size_t)nb2; off < n && st == DGX_OK; off += (size_t)1 << 28)
			if (dgx::launch_fill_u64(c.stream, reinterpret_cast<unsigned long long*>(values + off),
			                         (int)std::min(n - off, (size_t)1 << 28), 0ull) != hipSuccess)
				st = DGX_ERR_HIP;
		if (st != DGX_OK) break;
		// each block once: the own block, then the faces one local face at a time (a neighbour reached through two
		// faces, or the element itself on a periodic mesh one element wide, gets their sum)
		if (dgx::launch_bsr_place(c.stream, dblk, (size_t)nb2, 0, slots, values, E, nb2, false) != hipSuccess) { st = DGX_ERR_HIP; break; }
		for (int lf = 0; lf < nf && st == DGX_OK; ++lf)
			if (dgx::launch_bsr_place(c.stream, fblk, (size_t)nf * nb2, (size_t)lf * nb2, slots + (size_t)(lf + 1) * E, values,
			                          E, nb2, true) != hipSuccess)
				st = DGX_ERR_HIP;
	} while (false);
	if (hipStreamSynchronize(c.stream) != hipSuccess && st == DGX_OK) st = DGX_ERR_HIP;
	if (dblk) (void)hipFree(dblk);
	if (fblk) (void)hipFree(fblk);
	if (slots) (void)hipFree(slots);
	return st;
}

dgx_status dgx_cfl_dt(dgx_ctx* h, const double* u, double cfl, double* inv_dt)
{
	if (!h || !u || !inv_dt || !(cfl > 0.0)) return DGX_ERR_ARG;
	Ctx& c = h->c;
	if (!c.mesh_ready) return c.hp ? DGX_ERR_UNSUPPORTED : DGX_ERR_NO_MESH;
	if (c.nv != c.d + 2) return DGX_ERR_UNSUPPORTED;   // Euler and Navier-Stokes, 2-D and 3-D
	return hip_status(dgx::launch_cfl_dt(c, u, cfl, inv_dt));
}

dgx_status dgx_mass_sub(dgx_ctx* h, const double* x, const double* scale, double* y)
{
	if (!h || !x || !y) return DGX_ERR_ARG;
	Ctx& c = h->c;
	if (!c.mesh_ready) return c.hp ? DGX_ERR_UNSUPPORTED : DGX_ERR_NO_MESH;
	if (c.d != 2 && c.d != 3) return DGX_ERR_UNSUPPORTED;
	return hip_status(dgx::launch_mass_sub(c, x, scale, y));
}

dgx_status dgx_newton_operator(dgx_ctx* h, const double* u, const double* mass_scale, const double* v, double* out)
{
	if (!h || !u || !mass_scale || !v || !out) return DGX_ERR_ARG;
	Ctx& c = h->c;
	if (!c.mesh_ready) return c.hp ? DGX_ERR_UNSUPPORTED : DGX_ERR_NO_MESH;
	if (c.d != 2 && c.d != 3) return DGX_ERR_UNSUPPORTED;
	return hip_status(dgx::launch_jvp(c, u, v, out, mass_scale));
}

dgx_status dgx_jacobian_vector(dgx_ctx* h, const double* u, const double* v, double* jv)
{
	if (!h || !u || !v || !jv) return DGX_ERR_ARG;
	Ctx& c = h->c;
	if (!c.mesh_ready) return c.hp ? DGX_ERR_UNSUPPORTED : DGX_ERR_NO_MESH;
	if (c.d != 2 && c.d != 3) return DGX_ERR_UNSUPPORTED;
	return hip_status(dgx::launch_jvp(c, u, v, jv));
}

dgx_status dgx_block_inverse(dgx_ctx* h, double* blocks, int E, int nb)
{
	if (!h || !blocks || E < 0 || nb <= 0 || nb > dgx::kMaxInvBlock) return DGX_ERR_ARG;
	return hip_status(dgx::launch_block_inverse(h->c.stream, blocks, E, nb, nullptr));
}

dgx_status dgx_block_inv_apply(dgx_ctx* h, const double* inv, const double* x, double* y, int E, int nb)
{
	if (!h || !inv || !x || !y || E < 0 || nb <= 0 || nb > dgx::kMaxInvBlock) return DGX_ERR_ARG;
	return hip_status(dgx::launch_block_inv_apply(h->c.stream, inv, x, y, E, nb));
}

dgx_status dgx_block_lu_factor(dgx_ctx* h, double* blocks, int* piv, int E, int nb)
{
	if (!h || !blocks || !piv || E < 0 || nb <= 0 || nb > dgx::kMaxLuBlock) return DGX_ERR_ARG;  // nb^2 doubles in LDS
	return hip_status(dgx::launch_block_lu(h->c.stream, blocks, piv, E, nb));
}

dgx_status dgx_block_lu_solve(dgx_ctx* h, const double* lu, const int* piv, double* x, int E, int nb)
{
	if (!h || !lu || !piv || !x || E < 0 || nb <= 0 || nb > dgx::kMaxLuBlock) return DGX_ERR_ARG;
	return hip_status(dgx::launch_block_lu_solve(h->c.stream, lu, piv, x, E, nb));
}

}  // extern "C"
