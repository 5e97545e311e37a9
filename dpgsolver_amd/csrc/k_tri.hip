// k_tri.hip -- TRI elements of 2-D mixed meshes and the faces that touch them (FP64 Euler).
//
// TRI group residual (k_tri_rhs): the volume term of compute_volume_rlhs_dg with the element's dense operators
// (compute_volume_rlhs_T.c:145-177: s = cv0_vs_vc u, fr_d = sum_j m_{jd} f_j, rhs += sum_d tw1_vt_vc[d] fr_d), and
// the face term of every face read as the lifted flux z that k_mix_faces wrote at the element's own GL face
// points (rhs -= tw0_vt_fc z, compute_face_rlhs_T.c:297-359).  One element per wavefront, four per workgroup,
// the order's operators staged in LDS once per workgroup over a strided element loop.
//
// Faces with a TRI side (k_mix_faces): one face per thread.  Each side's trace at its own GL points (TRI: the
// dense trace of all its nodes; QUAD: its GLL face nodes), interpolated to the face's GL points (order
// max(p_L, p_R), compute_face_reference_order face_solver.c:123-131), the right side in the left ordering; the
// numerical flux with n J_f of the lower-order side (get_dominant_geom_vol_side_index, face_solver.c:95-107);
// then each side's z = P^T (w J_f F) / w_own at its own GL points (the right side negated, own ordering), which
// the element's own test functions lift exactly (they are polynomials of its order along the face).
#include "k_residual.h"

namespace dgx {

__device__ inline int quad_face_node(int N, int lf, int i)
{
	const int fd = lf >> 1, fix = (lf & 1) ? N - 1 : 0;
	return fd == 0 ? fix + N * i : i + N * fix;
}

template <int P>
constexpr int tri_nc()
{
	return P == 1 ? 3 : P == 2 ? 6 : P == 3 ? 12 : P == 4 ? 16 : 25;   // WV of degree 2p
}

struct TriRhsArgs {
	const double* u;     // the group's state, E x nv x Np
	double* r;           // the group's residual
	const double* met;   // E x 4 x Nc
	const double* z;     // the group's lifted face fluxes, E x 3 x nv x N
	const TriOps<double>* ops;
	int E;
};

template <int P>
__global__ void __launch_bounds__(256) k_tri_rhs(TriRhsArgs a)
{
	constexpr int Np = (P + 1) * (P + 2) / 2, Nc = tri_nc<P>(), N = P + 1, NV = 4, W = 4;
	__shared__ double sI[Nc * Np], sDrW[Np * Nc], sDsW[Np * Nc], sTw0[3 * Np * N];
	__shared__ double sU[W][NV * Np], sFR[W][2 * NV * Nc], sZ[W][3 * NV * N];
	for (int i = threadIdx.x; i < Nc * Np; i += blockDim.x) {
		sI[i] = a.ops->I[i];
		sDrW[i] = a.ops->DrW[i];
		sDsW[i] = a.ops->DsW[i];
	}
	for (int i = threadIdx.x; i < 3 * Np * N; i += blockDim.x) sTw0[i] = a.ops->tw0[i / (Np * N)][i % (Np * N)];
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
	for (int base = blockIdx.x * W; base < a.E; base += gridDim.x * W) {   // the same trip count for every wave
		const int e = base + wave;
		const bool valid = e < a.E;
		__syncthreads();   // the operators are staged / the previous element's buffers are read
		if (valid) {
			for (int i = lane; i < NV * Np; i += 64) sU[wave][i] = a.u[(size_t)e * NV * Np + i];
			for (int i = lane; i < 3 * NV * N; i += 64) sZ[wave][i] = a.z[(size_t)e * 3 * NV * N + i];
		}
		__syncthreads();
		if (valid && lane < Nc) {
			const int q = lane;
			double s[NV];
#pragma unroll
			for (int v = 0; v < NV; ++v) {
				double acc = 0.0;
#pragma unroll
				for (int m = 0; m < Np; ++m) acc += sI[q * Np + m] * sU[wave][v * Np + m];
				s[v] = acc;
			}
			const double* mp = a.met + (size_t)e * 4 * Nc + q;
			const double m0 = mp[0], m1 = mp[Nc], m2 = mp[2 * Nc], m3 = mp[3 * Nc];
			double f[2 * NV];
			euler_flux<double, 2>(s, f);
#pragma unroll
			for (int v = 0; v < NV; ++v) {
				sFR[wave][v * Nc + q] = m0 * f[2 * v] + m1 * f[2 * v + 1];
				sFR[wave][(NV + v) * Nc + q] = m2 * f[2 * v] + m3 * f[2 * v + 1];
			}
		}
		__syncthreads();
		if (valid && lane < Np) {
			const int m = lane;
#pragma unroll
			for (int v = 0; v < NV; ++v) {
				double acc = 0.0;
#pragma unroll
				for (int q = 0; q < Nc; ++q)
					acc += sDrW[m * Nc + q] * sFR[wave][v * Nc + q] + sDsW[m * Nc + q] * sFR[wave][(NV + v) * Nc + q];
#pragma unroll
				for (int lf = 0; lf < 3; ++lf)
#pragma unroll
					for (int k = 0; k < N; ++k) acc -= sTw0[(lf * Np + m) * N + k] * sZ[wave][(lf * NV + v) * N + k];
				a.r[(size_t)e * NV * Np + v * Np + m] = acc;
			}
		}
	}
}

hipError_t launch_tri_rhs64(const TriGroup& g, const TriOps<double>* ops, const double* u, double* r,
                            const double* z, hipStream_t st)
{
	if (g.E == 0) return hipSuccess;
	TriRhsArgs a{ u + g.u_off, r + g.u_off, g.met, z + g.z_off, ops + g.p, g.E };
	const int grid = std::min((g.E + 3) / 4, 4096);
	switch (g.p) {
	case 1: hipLaunchKernelGGL(k_tri_rhs<1>, dim3(grid), dim3(256), 0, st, a); break;
	case 2: hipLaunchKernelGGL(k_tri_rhs<2>, dim3(grid), dim3(256), 0, st, a); break;
	case 3: hipLaunchKernelGGL(k_tri_rhs<3>, dim3(grid), dim3(256), 0, st, a); break;
	case 4: hipLaunchKernelGGL(k_tri_rhs<4>, dim3(grid), dim3(256), 0, st, a); break;
	case 5: hipLaunchKernelGGL(k_tri_rhs<5>, dim3(grid), dim3(256), 0, st, a); break;
	default: return hipErrorInvalidValue;
	}
	return hipGetLastError();
}

struct MixFaceArgs {
	const MixFace* faces;
	const double* nJ;              // Fmix x kHpMaxN x 2
	const double* u;               // the whole state
	double* z;                     // the whole lifted-flux buffer
	const TriOps<double>* tops;    // indexed by p
	const HpOps<double>* hops;
	double bc_prim[5];
	int F;
};

// one side's trace at the face's Nf GL points, in that side's own face ordering
__device__ inline void side_trace(const MixFaceArgs& a, int kind, int Ns, int lf, long long uoff, int Nf,
                                  double (*out)[4])
{
	constexpr int NV = 4;
	const double* U = a.u + uoff;
	if (kind == 1) {
		const TriOps<double>& o = a.tops[Ns - 1];
		const int Np = o.Np;
		double t[kHpMaxN][NV];
		for (int k = 0; k < Ns; ++k)
#pragma unroll
			for (int v = 0; v < NV; ++v) {
				double acc = 0.0;
				for (int m = 0; m < Np; ++m) acc += o.Ifc[lf][k * Np + m] * U[v * Np + m];
				t[k][v] = acc;
			}
		const double* Pg = a.hops->Pg[Ns][Nf];   // own GL points -> face GL points
		for (int q = 0; q < Nf; ++q)
#pragma unroll
			for (int v = 0; v < NV; ++v) {
				double acc = 0.0;
				for (int k = 0; k < Ns; ++k) acc += Pg[q * Ns + k] * t[k][v];
				out[q][v] = acc;
			}
		return;
	}
	const double* Ig = a.hops->Ig[Ns][Nf];       // GLL face nodes -> face GL points
	const int Npq = Ns * Ns;
	for (int q = 0; q < Nf; ++q)
#pragma unroll
		for (int v = 0; v < NV; ++v) {
			double acc = 0.0;
			for (int i = 0; i < Ns; ++i) acc += Ig[q * Ns + i] * U[v * Npq + quad_face_node(Ns, lf, i)];
			out[q][v] = acc;
		}
}

template <int FLUX>
__global__ void __launch_bounds__(128) k_mix_faces(MixFaceArgs a)
{
	constexpr int NV = 4;
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= a.F) return;
	const MixFace f = a.faces[i];
	const bool inter = (f.bc == DGX_BC_INTERIOR);
	const int Nf = inter ? (f.NL > f.NR ? f.NL : f.NR) : f.NL;
	double sL[kHpMaxN][NV], sRo[kHpMaxN][NV], g[kHpMaxN][NV];
	side_trace(a, f.kL, f.NL, f.lfL, f.uL, Nf, sL);
	if (inter) side_trace(a, f.kR, f.NR, f.lfR, f.uR, Nf, sRo);
	const double* wf = a.hops->w[Nf];
	for (int q = 0; q < Nf; ++q) {
		const double* nj = a.nJ + ((size_t)i * kHpMaxN + q) * 2;
		const double Jf = sqrt(nj[0] * nj[0] + nj[1] * nj[1]);
		const double nu[2] = { nj[0] / Jf, nj[1] / Jf };
		double sR[NV], nnf[NV];
		if (inter) {
			const int qr = f.rev ? Nf - 1 - q : q;
#pragma unroll
			for (int v = 0; v < NV; ++v) sR[v] = sRo[qr][v];
		} else {
			bc_state<double, 2>(f.bc, sL[q], nu, a.bc_prim, sR);
		}
		if constexpr (FLUX == DGX_FLUX_LF) numflux_lf<double, 2>(sL[q], sR, nu, nnf);
		else numflux_roe<double, 2>(sL[q], sR, nu, nnf);
#pragma unroll
		for (int v = 0; v < NV; ++v) g[q][v] = wf[q] * Jf * nnf[v];
	}
	// lifts to each side's own GL face points: z = P^T g / w_own (the right side negated, in its own ordering)
	{
		const double* Pg = a.hops->Pg[f.NL][Nf];
		for (int k = 0; k < f.NL; ++k)
#pragma unroll
			for (int v = 0; v < NV; ++v) {
				double acc = 0.0;
				for (int q = 0; q < Nf; ++q) acc += Pg[q * f.NL + k] * g[q][v];
				a.z[f.zL + v * f.NL + k] = acc * a.hops->winv[f.NL][k];
			}
	}
	if (inter) {
		const double* Pg = a.hops->Pg[f.NR][Nf];
		for (int k = 0; k < f.NR; ++k)
#pragma unroll
			for (int v = 0; v < NV; ++v) {
				double acc = 0.0;
				for (int q = 0; q < Nf; ++q) acc += Pg[(f.rev ? Nf - 1 - q : q) * f.NR + k] * g[q][v];
				a.z[f.zR + v * f.NR + k] = 0.0 - acc * a.hops->winv[f.NR][k];
			}
	}
}

hipError_t launch_mix_faces64(const HpData& H, const Ctx& c, const double* u, hipStream_t st)
{
	if (H.Fmix == 0) return hipSuccess;
	MixFaceArgs a{};
	a.faces = H.mfaces;
	a.nJ = H.mnJ;
	a.u = u;
	a.z = H.z64;
	a.tops = H.tri_ops64;
	a.hops = H.ops64;
	for (int k = 0; k < 5; ++k) a.bc_prim[k] = c.cfg.bc_prim[k];
	a.F = H.Fmix;
	const int grid = (H.Fmix + 127) / 128;
	if (c.cfg.num_flux == DGX_FLUX_LF) hipLaunchKernelGGL(k_mix_faces<DGX_FLUX_LF>, dim3(grid), dim3(128), 0, st, a);
	else hipLaunchKernelGGL(k_mix_faces<DGX_FLUX_ROE>, dim3(grid), dim3(128), 0, st, a);
	return hipGetLastError();
}


// ------------------------------------------------------------------------------------------- Navier-Stokes (BR2)
// TRI groups with the BR2 viscous terms (compute_grad_coef_dg_T.c:171-307, compute_face_rlhs_dg_T.c:243-300):
//   k_tri_gv:  the volume part of the weak gradient G_v = M^-1 I^T W (J grad s) per element, and its traces at the
//              element's own GL face points (read by the face kernel);
//   k_mix_faces_ns: per face with a TRI side, q = n J_f (s_R - s_L) / 2 at the face's GL points, each side's
//              projection y = P^T w q / w_own at its own GL points (written for the residual kernels after the
//              lifted flux), its face gradient h = trace G_v + eta (lift trace) y with eta = 1.01 NFMAX = 4.04
//              (TRI: the dense K_f = Ifc_f M^-1 tw0_f; QUAD: its diagonal kappa), the inviscid flux plus the
//              central n . (fv_L + fv_R) / 2 (numerical_flux_central_T.c:76-104), lifted as for Euler;
//   k_tri_ns:  G = G_v + sum_f M^-1 tw0_f y_f at the cubature, F_euler + fv (flux_navier_stokes_T.c:1005-1040) in
//              the volume term, and the lifted face fluxes.
struct TriNsArgs {
	const double* u;     // the group's state
	double* r;           // the group's residual
	const double* met;
	const double* minv;  // E x Np x Np
	double* gv;          // E x 2 x nv x Np
	double* gvt;         // E x 3 x 2 x nv x N
	const double* z;     // E x 3 x (3 nv N): lifted flux [v][a], then y [d][v][a]
	const TriOps<double>* ops;
	double mu, Pr;
	int E;
};

template <int P>
__global__ void __launch_bounds__(256) k_tri_gv(TriNsArgs a)
{
	constexpr int Np = (P + 1) * (P + 2) / 2, Nc = tri_nc<P>(), N = P + 1, NV = 4, W = 4;
	__shared__ double sI[Nc * Np], sDr[Nc * Np], sDs[Nc * Np], sIfc[3 * N * Np], sWq[Nc];
	__shared__ double sU[W][NV * Np], sX[W][2 * NV * Nc], sB[W][2 * NV * Np];
	for (int i = threadIdx.x; i < Nc * Np; i += blockDim.x) {
		sI[i] = a.ops->I[i];
		sDr[i] = a.ops->Dr[i];
		sDs[i] = a.ops->Ds[i];
	}
	for (int i = threadIdx.x; i < 3 * N * Np; i += blockDim.x) sIfc[i] = a.ops->Ifc[i / (N * Np)][i % (N * Np)];
	for (int i = threadIdx.x; i < Nc; i += blockDim.x) sWq[i] = a.ops->w[i];
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
	for (int base = blockIdx.x * W; base < a.E; base += gridDim.x * W) {
		const int e = base + wave;
		const bool valid = e < a.E;
		__syncthreads();
		if (valid)
			for (int i = lane; i < NV * Np; i += 64) sU[wave][i] = a.u[(size_t)e * NV * Np + i];
		__syncthreads();
		if (valid && lane < Nc) {   // W J grad s at the cubature
			const int q = lane;
			const double* mp = a.met + (size_t)e * 4 * Nc + q;
			const double m0 = mp[0], m1 = mp[Nc], m2 = mp[2 * Nc], m3 = mp[3 * Nc];
#pragma unroll
			for (int v = 0; v < NV; ++v) {
				double sr = 0.0, ss = 0.0;
#pragma unroll
				for (int m = 0; m < Np; ++m) {
					sr += sDr[q * Np + m] * sU[wave][v * Np + m];
					ss += sDs[q * Np + m] * sU[wave][v * Np + m];
				}
				sX[wave][v * Nc + q] = sWq[q] * (m0 * sr + m2 * ss);
				sX[wave][(NV + v) * Nc + q] = sWq[q] * (m1 * sr + m3 * ss);
			}
		}
		__syncthreads();
		if (valid && lane < Np) {   // I^T (W J grad s)
			const int m = lane;
#pragma unroll
			for (int k = 0; k < 2 * NV; ++k) {
				double acc = 0.0;
#pragma unroll
				for (int q = 0; q < Nc; ++q) acc += sI[q * Np + m] * sX[wave][k * Nc + q];
				sB[wave][k * Np + m] = acc;
			}
		}
		__syncthreads();
		double g[2 * NV];
		if (valid && lane < Np) {   // G_v = M^-1 (...)
			const int m = lane;
			const double* mi = a.minv + (size_t)e * Np * Np + m * Np;
#pragma unroll
			for (int k = 0; k < 2 * NV; ++k) g[k] = 0.0;
			for (int n = 0; n < Np; ++n) {
				const double c = mi[n];
#pragma unroll
				for (int k = 0; k < 2 * NV; ++k) g[k] += c * sB[wave][k * Np + n];
			}
		}
		__syncthreads();
		if (valid && lane < Np) {
#pragma unroll
			for (int k = 0; k < 2 * NV; ++k) {
				sB[wave][k * Np + lane] = g[k];
				a.gv[((size_t)e * 2 * NV + k) * Np + lane] = g[k];
			}
		}
		__syncthreads();
		if (valid && lane < 3 * N) {   // traces at the own GL face points
			const int lf = lane / N, k0 = lane - lf * N;
#pragma unroll
			for (int k = 0; k < 2 * NV; ++k) {
				double acc = 0.0;
#pragma unroll
				for (int m = 0; m < Np; ++m) acc += sIfc[(lf * N + k0) * Np + m] * sB[wave][k * Np + m];
				a.gvt[(((size_t)e * 3 + lf) * 2 * NV + k) * N + k0] = acc;
			}
		}
	}
}

template <int P>
__global__ void __launch_bounds__(256) k_tri_ns(TriNsArgs a)
{
	constexpr int Np = (P + 1) * (P + 2) / 2, Nc = tri_nc<P>(), N = P + 1, NV = 4, W = 4, ZB = 3 * NV * N;
	__shared__ double sI[Nc * Np], sDrW[Np * Nc], sDsW[Np * Nc], sTw0[3 * Np * N];
	__shared__ double sU[W][NV * Np], sFR[W][2 * NV * Nc], sZ[W][3 * ZB], sG[W][2 * NV * Np];
	for (int i = threadIdx.x; i < Nc * Np; i += blockDim.x) {
		sI[i] = a.ops->I[i];
		sDrW[i] = a.ops->DrW[i];
		sDsW[i] = a.ops->DsW[i];
	}
	for (int i = threadIdx.x; i < 3 * Np * N; i += blockDim.x) sTw0[i] = a.ops->tw0[i / (Np * N)][i % (Np * N)];
	const double mu = a.mu, Pr = a.Pr;
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
	for (int base = blockIdx.x * W; base < a.E; base += gridDim.x * W) {
		const int e = base + wave;
		const bool valid = e < a.E;
		__syncthreads();
		if (valid) {
			for (int i = lane; i < NV * Np; i += 64) sU[wave][i] = a.u[(size_t)e * NV * Np + i];
			for (int i = lane; i < 3 * ZB; i += 64) sZ[wave][i] = a.z[(size_t)e * 3 * ZB + i];
		}
		__syncthreads();
		// the face lifts of the gradient: t = sum_f tw0_f y_f at the nodes (into sFR, free until the fluxes)
		if (valid && lane < Np) {
			const int m = lane;
#pragma unroll
			for (int k = 0; k < 2 * NV; ++k) {
				double acc = 0.0;
#pragma unroll
				for (int lf = 0; lf < 3; ++lf)
#pragma unroll
					for (int b = 0; b < N; ++b) acc += sTw0[(lf * Np + m) * N + b] * sZ[wave][lf * ZB + NV * N + k * N + b];
				sFR[wave][k * Np + m] = acc;
			}
		}
		__syncthreads();
		if (valid && lane < Np) {   // G = G_v + M^-1 t
			const int m = lane;
			const double* mi = a.minv + (size_t)e * Np * Np + m * Np;
			double g[2 * NV];
#pragma unroll
			for (int k = 0; k < 2 * NV; ++k) g[k] = a.gv[((size_t)e * 2 * NV + k) * Np + m];
			for (int n = 0; n < Np; ++n) {
				const double c = mi[n];
#pragma unroll
				for (int k = 0; k < 2 * NV; ++k) g[k] += c * sFR[wave][k * Np + n];
			}
#pragma unroll
			for (int k = 0; k < 2 * NV; ++k) sG[wave][k * Np + m] = g[k];
		}
		__syncthreads();
		if (valid && lane < Nc) {
			const int q = lane;
			double s[NV], gq[2 * NV];
#pragma unroll
			for (int v = 0; v < NV; ++v) {
				double acc = 0.0;
#pragma unroll
				for (int m = 0; m < Np; ++m) acc += sI[q * Np + m] * sU[wave][v * Np + m];
				s[v] = acc;
			}
#pragma unroll
			for (int k = 0; k < 2 * NV; ++k) {
				double acc = 0.0;
#pragma unroll
				for (int m = 0; m < Np; ++m) acc += sI[q * Np + m] * sG[wave][k * Np + m];
				gq[k] = acc;
			}
			const double* mp = a.met + (size_t)e * 4 * Nc + q;
			const double m0 = mp[0], m1 = mp[Nc], m2 = mp[2 * Nc], m3 = mp[3 * Nc];
			double f[2 * NV], fv[2 * NV];
			euler_flux<double, 2>(s, f);
			ns_viscous_flux2<double>(s, gq, mu, Pr, fv);
#pragma unroll
			for (int k = 0; k < 2 * NV; ++k) f[k] += fv[k];
#pragma unroll
			for (int v = 0; v < NV; ++v) {
				sFR[wave][v * Nc + q] = m0 * f[2 * v] + m1 * f[2 * v + 1];
				sFR[wave][(NV + v) * Nc + q] = m2 * f[2 * v] + m3 * f[2 * v + 1];
			}
		}
		__syncthreads();
		if (valid && lane < Np) {
			const int m = lane;
#pragma unroll
			for (int v = 0; v < NV; ++v) {
				double acc = 0.0;
#pragma unroll
				for (int q = 0; q < Nc; ++q)
					acc += sDrW[m * Nc + q] * sFR[wave][v * Nc + q] + sDsW[m * Nc + q] * sFR[wave][(NV + v) * Nc + q];
#pragma unroll
				for (int lf = 0; lf < 3; ++lf)
#pragma unroll
					for (int b = 0; b < N; ++b) acc -= sTw0[(lf * Np + m) * N + b] * sZ[wave][lf * ZB + v * N + b];
				a.r[(size_t)e * NV * Np + v * Np + m] = acc;
			}
		}
	}
}

struct MixFaceNsArgs {
	MixFaceArgs base;
	const double* qgvt[kHpMaxN + 1];   // QUAD groups by N: face traces of G_v (E_g x 4 x 2 x nv x N) and kappa
	const double* qkap[kHpMaxN + 1];
	const double* tgvt[kTriMaxP];      // TRI groups by group index: gvt and K
	const double* tK[kTriMaxP];
	double mu, Pr;
};

// one side's face gradient h at its own GL points from its projection y: trace G_v + eta (lift trace) y
__device__ inline void side_face_gradient(const MixFaceNsArgs& a, int kind, int Ns, int lf, int el, int tg,
                                          const double (*y)[8], double (*h)[8])
{
	constexpr int NV = 4;
	const double eta = 1.01 * 4.0;
	if (kind == 1) {
		const double* t = a.tgvt[tg] + ((size_t)el * 3 + lf) * 2 * NV * Ns;
		const double* K = a.tK[tg] + ((size_t)el * 3 + lf) * Ns * Ns;
		for (int k0 = 0; k0 < Ns; ++k0)
			for (int k = 0; k < 2 * NV; ++k) {
				double acc = 0.0;
				for (int b = 0; b < Ns; ++b) acc += K[k0 * Ns + b] * y[b][k];
				h[k0][k] = t[k * Ns + k0] + eta * acc;
			}
		return;
	}
	const double* t = a.qgvt[Ns] + ((size_t)el * 4 + lf) * 2 * NV * Ns;
	const double* kap = a.qkap[Ns] + ((size_t)el * 4 + lf) * Ns;
	for (int k0 = 0; k0 < Ns; ++k0)
		for (int k = 0; k < 2 * NV; ++k) h[k0][k] = t[k * Ns + k0] + eta * kap[k0] * y[k0][k];
}

template <int FLUX>
__global__ void __launch_bounds__(128) k_mix_faces_ns(MixFaceNsArgs A)
{
	constexpr int NV = 4;
	const MixFaceArgs& a = A.base;
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= a.F) return;
	const MixFace f = a.faces[i];
	const int Nf = f.NL > f.NR ? f.NL : f.NR;
	double sL[kHpMaxN][NV], sRo[kHpMaxN][NV], qv[kHpMaxN][2 * NV];
	side_trace(a, f.kL, f.NL, f.lfL, f.uL, Nf, sL);
	side_trace(a, f.kR, f.NR, f.lfR, f.uR, Nf, sRo);
	const double* wf = a.hops->w[Nf];
	for (int q = 0; q < Nf; ++q) {   // w q, q = n J_f (s_R - s_L) / 2, left ordering
		const double* nj = a.nJ + ((size_t)i * kHpMaxN + q) * 2;
		const int qr = f.rev ? Nf - 1 - q : q;
#pragma unroll
		for (int v = 0; v < NV; ++v) {
			const double dl = 0.5 * (sRo[qr][v] - sL[q][v]);
			qv[q][v] = wf[q] * (nj[0] * dl);
			qv[q][NV + v] = wf[q] * (nj[1] * dl);
		}
	}
	// projections to each side's own points (written after the lifted flux), and the face gradients there
	double yL[kHpMaxN][8], yR[kHpMaxN][8], hL[kHpMaxN][8], hR[kHpMaxN][8];
	{
		const double* Pg = a.hops->Pg[f.NL][Nf];
		for (int k0 = 0; k0 < f.NL; ++k0)
			for (int k = 0; k < 2 * NV; ++k) {
				double acc = 0.0;
				for (int q = 0; q < Nf; ++q) acc += Pg[q * f.NL + k0] * qv[q][k];
				yL[k0][k] = acc * a.hops->winv[f.NL][k0];
				a.z[f.zL + NV * f.NL + k * f.NL + k0] = yL[k0][k];
			}
		const double* PgR = a.hops->Pg[f.NR][Nf];
		for (int k0 = 0; k0 < f.NR; ++k0)
			for (int k = 0; k < 2 * NV; ++k) {
				double acc = 0.0;
				for (int q = 0; q < Nf; ++q) acc += PgR[(f.rev ? Nf - 1 - q : q) * f.NR + k0] * qv[q][k];
				yR[k0][k] = acc * a.hops->winv[f.NR][k0];
				a.z[f.zR + NV * f.NR + k * f.NR + k0] = yR[k0][k];
			}
	}
	side_face_gradient(A, f.kL, f.NL, f.lfL, f.eL, f.tL, yL, hL);
	side_face_gradient(A, f.kR, f.NR, f.lfR, f.eR, f.tR, yR, hR);
	double g[kHpMaxN][NV];
	const double mu = A.mu, Pr = A.Pr;
	const double* PgL = a.hops->Pg[f.NL][Nf];
	const double* PgR = a.hops->Pg[f.NR][Nf];
	for (int q = 0; q < Nf; ++q) {
		const int qr = f.rev ? Nf - 1 - q : q;
		double gL[2 * NV], gR[2 * NV];
		for (int k = 0; k < 2 * NV; ++k) {
			double aL = 0.0, aR = 0.0;
			for (int k0 = 0; k0 < f.NL; ++k0) aL += PgL[q * f.NL + k0] * hL[k0][k];
			for (int k0 = 0; k0 < f.NR; ++k0) aR += PgR[qr * f.NR + k0] * hR[k0][k];
			gL[k] = aL;
			gR[k] = aR;
		}
		const double* nj = a.nJ + ((size_t)i * kHpMaxN + q) * 2;
		const double Jf = sqrt(nj[0] * nj[0] + nj[1] * nj[1]);
		const double nu[2] = { nj[0] / Jf, nj[1] / Jf };
		double sR[NV], nnf[NV], fvL[2 * NV], fvR[2 * NV];
#pragma unroll
		for (int v = 0; v < NV; ++v) sR[v] = sRo[qr][v];
		if constexpr (FLUX == DGX_FLUX_LF) numflux_lf<double, 2>(sL[q], sR, nu, nnf);
		else numflux_roe<double, 2>(sL[q], sR, nu, nnf);
		ns_viscous_flux2<double>(sL[q], gL, mu, Pr, fvL);
		ns_viscous_flux2<double>(sR, gR, mu, Pr, fvR);
#pragma unroll
		for (int v = 0; v < NV; ++v) {
			const double vis = 0.5 * ((fvL[2 * v] + fvR[2 * v]) * nu[0] + (fvL[2 * v + 1] + fvR[2 * v + 1]) * nu[1]);
			g[q][v] = wf[q] * Jf * (nnf[v] + vis);
		}
	}
	{
		const double* Pg = a.hops->Pg[f.NL][Nf];
		for (int k = 0; k < f.NL; ++k)
#pragma unroll
			for (int v = 0; v < NV; ++v) {
				double acc = 0.0;
				for (int q = 0; q < Nf; ++q) acc += Pg[q * f.NL + k] * g[q][v];
				a.z[f.zL + v * f.NL + k] = acc * a.hops->winv[f.NL][k];
			}
		const double* Pr_ = a.hops->Pg[f.NR][Nf];
		for (int k = 0; k < f.NR; ++k)
#pragma unroll
			for (int v = 0; v < NV; ++v) {
				double acc = 0.0;
				for (int q = 0; q < Nf; ++q) acc += Pr_[(f.rev ? Nf - 1 - q : q) * f.NR + k] * g[q][v];
				a.z[f.zR + v * f.NR + k] = 0.0 - acc * a.hops->winv[f.NR][k];
			}
	}
}

hipError_t launch_tri_ns_phase64(const TriGroup& g, const TriOps<double>* ops, const double* u, double* r,
                                 const double* z, double mu, double Pr, int phase, hipStream_t st)
{
	if (g.E == 0) return hipSuccess;
	TriNsArgs a{ u + g.u_off, r + g.u_off, g.met, g.minv, g.gv, g.gvt, z + g.z_off, ops + g.p, mu, Pr, g.E };
	const int grid = std::min((g.E + 3) / 4, 4096);
#define DGX_TRI_NS(PP)                                                                                     \
	case PP:                                                                                               \
		if (phase == 1) hipLaunchKernelGGL(k_tri_gv<PP>, dim3(grid), dim3(256), 0, st, a);                  \
		else hipLaunchKernelGGL(k_tri_ns<PP>, dim3(grid), dim3(256), 0, st, a);                             \
		break;
	switch (g.p) {
		DGX_TRI_NS(1) DGX_TRI_NS(2) DGX_TRI_NS(3) DGX_TRI_NS(4) DGX_TRI_NS(5)
	default: return hipErrorInvalidValue;
	}
#undef DGX_TRI_NS
	return hipGetLastError();
}

hipError_t launch_mix_faces_ns64(const HpData& H, const Ctx& c, const double* u, hipStream_t st)
{
	if (H.Fmix == 0) return hipSuccess;
	MixFaceNsArgs A{};
	MixFaceArgs& a = A.base;
	a.faces = H.mfaces;
	a.nJ = H.mnJ;
	a.u = u;
	a.z = H.z64;
	a.tops = H.tri_ops64;
	a.hops = H.ops64;
	a.F = H.Fmix;
	for (size_t g = 0; g < H.groups.size(); ++g) {
		const Ctx& s = H.groups[g]->c;
		A.qgvt[s.N] = s.ns_gvt;
		A.qkap[s.N] = s.ns_kappa;
	}
	for (size_t t = 0; t < H.tri.size() && t < (size_t)kTriMaxP; ++t) {
		A.tgvt[t] = H.tri[t].gvt;
		A.tK[t] = H.tri[t].K;
	}
	A.mu = c.cfg.mu;
	A.Pr = c.cfg.Pr;
	const int grid = (H.Fmix + 127) / 128;
	if (c.cfg.num_flux == DGX_FLUX_LF) hipLaunchKernelGGL(k_mix_faces_ns<DGX_FLUX_LF>, dim3(grid), dim3(128), 0, st, A);
	else hipLaunchKernelGGL(k_mix_faces_ns<DGX_FLUX_ROE>, dim3(grid), dim3(128), 0, st, A);
	return hipGetLastError();
}

}  // namespace dgx
