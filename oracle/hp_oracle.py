"""Variable-order (per-element p) 2-D DG residual on the CPU, in numpy -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline may import this module; the product path
never does.  It restates the reference's per-volume / per-face residual loops for volumes of different
orders:

* volume term (compute_volume_rlhs_dg_T.c:53-106, compute_volume_rlhs_T.c:166-177): the element's own GLL
  solution basis of order p_e, interpolated to its (p_e+1)^2 GL cubature nodes, R += sum_d tw1_d fr_d with
  the J-scaled cofactor metrics of the element's own isoparametric map;
* face term (compute_face_rlhs_dg_T.c:88-202, compute_face_rlhs_T.c:297-359): the face order is the larger
  order of its two volumes (face_solver.c:123-131, compute_face_reference_order), so both traces are
  interpolated to the face's p_f+1 GL points (get_operator__cv0_vs_fc_T with p_f, compute_face_rlhs_T.c:67-80),
  the numerical flux is scaled by w J_f, and each side is lifted with its own test functions at those points
  (get_operator__tw0_vt_fc_T, compute_face_rlhs_T.c:51-64); the right side is permuted into the left
  ordering and negated (finalize_face_rhs_dg_like_T, compute_face_rlhs_T.c:347-359);
* face geometry: the normal and face Jacobian are taken from the lower-order side
  (get_dominant_geom_vol_side_index, face_solver.c:95-107; left on a tie), as n J_f = the outward rotation
  of the tangent of that side's face interpolant.

Navier-Stokes (rhs_ns, viscous=(mu, Pr)): the BR2 gradient of or_ns_gradient / or_compute_rhs_ns with the face
lifts at the face's GL points and each side's own test functions, and the central viscous face flux at the same
points from the traces of gv + 1.01 * 4 * gf[side].

Pinned by tests/test_hp_oracle.py: equal to the dense-operator C oracle (or_compute_rhs_d) on uniform-order
meshes, curved and with rotated elements; free-stream preservation and discrete conservation on
variable-order meshes (the first fails for a transposed or mis-weighted projection at a nonconforming face,
the second for a lift on the wrong side or a missing permutation); and convergence of M^-1 R to the exact
-div F of a smooth state on variable-order meshes.
"""
from __future__ import annotations

import numpy as np

import pyoracle as O

GAMMA, GM1 = 1.4, 0.4


class _Ops:
    """1-D operators of one order: GLL solution nodes, GL cubature of the same count."""

    def __init__(self, N):
        self.N = N
        self.xl, self.wl = O.gauss_lobatto(N)
        self.xg, self.wg = O.gauss_legendre(N)
        self.I, self.D = O.lagrange(self.xl, self.xg)        # GLL basis and its derivative at the GL nodes


_ops_cache = {}


def ops(N):
    if N not in _ops_cache:
        _ops_cache[N] = _Ops(N)
    return _ops_cache[N]


def face_nodes(N, lf):
    """Volume node index of each face node of local face lf (face coordinate = the other dimension)."""
    fd, fix = lf // 2, (N - 1 if lf % 2 else 0)
    i = np.arange(N)
    return fix + N * i if fd == 0 else i + N * fix


def outward_nJ(xf, N, lf, s):
    """n J_f of local face lf at face coordinates s, from the face interpolant of the face nodes xf (2, N)."""
    _, dl = O.lagrange(ops(N).xl, np.asarray(s, float))
    t = xf @ dl.T                                      # (2, nq): d x / d s along the face coordinate
    fd, side = lf // 2, lf % 2
    if fd == 0:                                        # xi = +-1, s = eta: J grad xi = (y_eta, -x_eta)
        n = np.stack([t[1], -t[0]])
    else:                                              # eta = +-1, s = xi: J grad eta = (-y_xi, x_xi)
        n = np.stack([-t[1], t[0]])
    return n if side else -n


class HpOracle:
    """One variable-order 2-D discretisation (Euler; mesh.HpMesh) on the CPU."""

    def __init__(self, mesh, pde, flux, bc_prim=None, viscous=None):
        assert pde == 1, "Euler (with viscous=(mu, Pr): Navier-Stokes through rhs_ns)"
        self.mesh, self.pde, self.flux = mesh, pde, flux
        self.viscous = viscous
        self.nv = 4
        self.bc_prim = np.zeros(5) if bc_prim is None else np.resize(np.asarray(bc_prim, float), 5)
        self.off = mesh.offsets(self.nv)
        self._geo = [self._volume_geometry(e) for e in range(mesh.E)]

    def _volume_geometry(self, e):
        N = int(self.mesh.p[e]) + 1
        o = ops(N)
        x = self.mesh.xyz[e]                                 # (2, Np)
        Dxi, Deta = np.kron(o.I, o.D), np.kron(o.D, o.I)     # r-fastest: node = i + N j
        xr, xs = x @ Dxi.T, x @ Deta.T                       # (2, Nc) derivatives at the GL nodes
        W = np.kron(o.wg, o.wg)
        J = xr[0] * xs[1] - xr[1] * xs[0]
        return dict(N=N, Dxi=Dxi, Deta=Deta, Ivol=np.kron(o.I, o.I), W=W, J=J, xr=xr, xs=xs)

    def element(self, u, e):
        g = self._geo[e]
        return u[self.off[e]:self.off[e + 1]].reshape(self.nv, g["N"] ** 2)

    def rhs(self, u):
        u = np.asarray(u, float)
        r = np.zeros_like(u)
        m = self.mesh
        for e in range(m.E):
            g = self._geo[e]
            U = self.element(u, e)
            s = (g["Ivol"] @ U.T)                            # (Nc, nv) at the GL nodes
            f = np.stack([O.euler_flux(2, s[q]).reshape(self.nv, 2) for q in range(s.shape[0])])  # (Nc, nv, 2)
            xr, xs = g["xr"], g["xs"]
            # fr_xi = y_eta f_x - x_eta f_y ; fr_eta = -y_xi f_x + x_xi f_y   (J-scaled cofactors)
            fr0 = xs[1][:, None] * f[:, :, 0] - xs[0][:, None] * f[:, :, 1]
            fr1 = -xr[1][:, None] * f[:, :, 0] + xr[0][:, None] * f[:, :, 1]
            R = g["Dxi"].T @ (g["W"][:, None] * fr0) + g["Deta"].T @ (g["W"][:, None] * fr1)   # (Np, nv)
            r[self.off[e]:self.off[e + 1]] += R.T.ravel()
        for f_ in range(m.F):
            self._face(u, r, m.faces[f_])
        return r

    def _face(self, u, r, row):
        eL, fL, eR, fR, ord_, bc = (int(v) for v in row)
        m = self.mesh
        NL = int(m.p[eL]) + 1
        NR = NL if bc != 0 else int(m.p[eR]) + 1
        Nf = max(NL, NR)
        xq, wq = O.gauss_legendre(Nf)
        perm = (Nf - 1 - np.arange(Nf)) if ord_ == 1 else np.arange(Nf)   # right index of left point q
        IL, _ = O.lagrange(ops(NL).xl, xq)                   # (Nf, NL): left face basis at the face points
        UL = self.element(u, eL)[:, face_nodes(NL, fL)]      # (nv, NL)
        sL = (IL @ UL.T)                                     # (Nf, nv) left ordering
        # face geometry from the lower-order side (left on a tie), in the left ordering, left outward
        if bc != 0 or NL <= NR:
            nJ = outward_nJ(m.xyz[eL][:, face_nodes(NL, fL)], NL, fL, xq)            # (2, Nf)
        else:
            nR = outward_nJ(m.xyz[eR][:, face_nodes(NR, fR)], NR, fR, xq)
            nJ = -nR[:, perm]
        Jf = np.sqrt((nJ ** 2).sum(0))
        n = nJ / Jf
        if bc == 0:
            IR, _ = O.lagrange(ops(NR).xl, xq)
            UR = self.element(u, eR)[:, face_nodes(NR, fR)]
            sR_own = IR @ UR.T                               # (Nf, nv) right ordering
            sR = sR_own[perm]
        else:
            sR = np.stack([O.bc_state(2, bc, sL[q], n[:, q], self.bc_prim) for q in range(Nf)])
        nnf = np.stack([O.numflux(2, self.pde, self.flux, sL[q], sR[q], n[:, q]) for q in range(Nf)])
        gq = (wq * Jf)[:, None] * nnf                        # (Nf, nv)
        # lift: left test functions at the left points, right test functions at the same physical points
        RL = IL.T @ gq                                       # (NL, nv)
        bL = self.off[eL]
        NpL = NL * NL
        for v in range(self.nv):
            r[bL + v * NpL + face_nodes(NL, fL)] -= RL[:, v]
        if bc == 0:
            gR = np.empty_like(gq)
            gR[perm] = gq                                    # right ordering
            RR = IR.T @ gR
            bR = self.off[eR]
            NpR = NR * NR
            for v in range(self.nv):
                r[bR + v * NpR + face_nodes(NR, fR)] += RR[:, v]

    def mass(self, e):
        g = self._geo[e]
        return g["Ivol"].T @ ((g["W"] * g["J"])[:, None] * g["Ivol"])

    def apply_minv(self, r):
        out = np.array(r, float)
        for e in range(self.mesh.E):
            Np = self._geo[e]["N"] ** 2
            R = out[self.off[e]:self.off[e + 1]].reshape(self.nv, Np)
            out[self.off[e]:self.off[e + 1]] = np.linalg.solve(self.mass(e), R.T).T.ravel()
        return out

    # ------------------------------------------------------------------------------------------------ BR2
    def _face_setup(self, u, row):
        """Traces and geometry of one interior face at its GL points (left ordering)."""
        eL, fL, eR, fR, ord_, bc = (int(v) for v in row)
        m = self.mesh
        NL, NR = int(m.p[eL]) + 1, int(m.p[eR]) + 1
        Nf = max(NL, NR)
        xq, wq = O.gauss_legendre(Nf)
        perm = (Nf - 1 - np.arange(Nf)) if ord_ == 1 else np.arange(Nf)
        IL, _ = O.lagrange(ops(NL).xl, xq)
        IR, _ = O.lagrange(ops(NR).xl, xq)
        if NL <= NR:
            nJ = outward_nJ(m.xyz[eL][:, face_nodes(NL, fL)], NL, fL, xq)
        else:
            nJ = -outward_nJ(m.xyz[eR][:, face_nodes(NR, fR)], NR, fR, xq)[:, perm]
        sL = IL @ self.element(u, eL)[:, face_nodes(NL, fL)].T
        sR = (IR @ self.element(u, eR)[:, face_nodes(NR, fR)].T)[perm]
        return dict(eL=eL, fL=fL, eR=eR, fR=fR, NL=NL, NR=NR, Nf=Nf, wq=wq, perm=perm, IL=IL, IR=IR, nJ=nJ,
                    sL=sL, sR=sR)

    def _face_vec(self, e, lf, N, I, vals):
        """sum_q phi_i(q) vals[q] onto the face nodes of element e: (Np,) per column of vals (nq, k)."""
        out = np.zeros((N * N, vals.shape[1]))
        out[face_nodes(N, lf)] = I.T @ vals
        return out

    def ns_gradient(self, u):
        """(g, gv, gf): weak gradient coefficients per element (d, nv, Np_e) -- volume part gv, full g -- and the
        per-face lifts gf[f] = (left, right) (compute_grad_coef_dg_T.c:171-307, central s_num)."""
        m = self.mesh
        gv = []
        for e in range(m.E):
            g = self._geo[e]
            U = self.element(u, e)
            sxi, seta = U @ g["Dxi"].T, U @ g["Deta"].T                      # (nv, Nc)
            xr, xs = g["xr"], g["xs"]
            Jg = [xs[1] * sxi - xr[1] * seta, -xs[0] * sxi + xr[0] * seta]   # J ds/dx, J ds/dy
            Minv = np.linalg.inv(self.mass(e))
            gv.append(np.stack([Minv @ (g["Ivol"].T @ (g["W"][:, None] * Jd.T)) for Jd in Jg]).transpose(0, 2, 1))
        gfull = [x.copy() for x in gv]
        gf = []
        for row in m.faces:
            fs = self._face_setup(u, row)
            dl = 0.5 * (fs["sR"] - fs["sL"])                                 # (Nf, nv)
            pair = []
            for side in (0, 1):
                e, lf, N, I = ((fs["eL"], fs["fL"], fs["NL"], fs["IL"]) if side == 0
                               else (fs["eR"], fs["fR"], fs["NR"], fs["IR"]))
                Minv = np.linalg.inv(self.mass(e))
                lifts = []
                for dim in range(2):
                    qv = (fs["wq"] * fs["nJ"][dim])[:, None] * dl           # w_q q_dim, left ordering
                    if side:
                        qr = np.empty_like(qv)
                        qr[fs["perm"]] = qv
                        qv = qr
                    lifts.append((Minv @ self._face_vec(e, lf, N, I, qv)).T)   # (nv, Np)
                lift = np.stack(lifts)
                gfull[e] += lift
                pair.append(lift)
            gf.append(pair)
        return gfull, gv, gf

    def rhs_ns(self, u):
        """Navier-Stokes (BR2) residual, no M^-1: F_euler + fv at the volume cubature from the full gradient;
        faces: nnf_euler + n . (fv_L + fv_R) / 2 with g_side = trace(gv + 1.01 * 4 * gf[side]) at the face's
        GL points (compute_face_rlhs_dg_T.c:243-300, numerical_flux_central_T.c:76-104)."""
        mu, Pr = self.viscous
        eta = 1.01 * 4
        u = np.asarray(u, float)
        m = self.mesh
        gfull, gv, gf = self.ns_gradient(u)
        r = np.zeros_like(u)
        for e in range(m.E):
            g = self._geo[e]
            U = self.element(u, e)
            s = g["Ivol"] @ U.T                                                 # (Nc, nv)
            gq = np.einsum("qi,dvi->qdv", g["Ivol"], gfull[e])                  # (Nc, d, nv)
            f = np.stack([O.euler_flux(2, s[q]).reshape(4, 2) + O.viscous_flux(2, s[q], gq[q].reshape(-1), mu, Pr)
                          for q in range(s.shape[0])])                 # gq[q].reshape(-1): g[v + 4 d]
            xr, xs = g["xr"], g["xs"]
            fr0 = xs[1][:, None] * f[:, :, 0] - xs[0][:, None] * f[:, :, 1]
            fr1 = -xr[1][:, None] * f[:, :, 0] + xr[0][:, None] * f[:, :, 1]
            R = g["Dxi"].T @ (g["W"][:, None] * fr0) + g["Deta"].T @ (g["W"][:, None] * fr1)
            r[self.off[e]:self.off[e + 1]] += R.T.ravel()
        for k, row in enumerate(m.faces):
            fs = self._face_setup(u, row)
            Nf = fs["Nf"]
            cL = gv[fs["eL"]] + eta * gf[k][0]
            cR = gv[fs["eR"]] + eta * gf[k][1]
            gL = np.einsum("qi,dvi->qdv", fs["IL"], cL[:, :, face_nodes(fs["NL"], fs["fL"])])
            gR = np.einsum("qi,dvi->qdv", fs["IR"], cR[:, :, face_nodes(fs["NR"], fs["fR"])])[fs["perm"]]
            Jf = np.sqrt((fs["nJ"] ** 2).sum(0))
            n = fs["nJ"] / Jf
            tot = np.empty((Nf, 4))
            for q in range(Nf):
                inv = O.numflux(2, self.pde, self.flux, fs["sL"][q], fs["sR"][q], n[:, q])
                fvL = O.viscous_flux(2, fs["sL"][q], gL[q].reshape(-1), mu, Pr)
                fvR = O.viscous_flux(2, fs["sR"][q], gR[q].reshape(-1), mu, Pr)
                tot[q] = inv + 0.5 * ((fvL + fvR) @ n[:, q])
            gq = (fs["wq"] * Jf)[:, None] * tot
            NL, NR = fs["NL"], fs["NR"]
            RL = fs["IL"].T @ gq
            for v in range(4):
                r[self.off[fs["eL"]] + v * NL * NL + face_nodes(NL, fs["fL"])] -= RL[:, v]
            gR_ = np.empty_like(gq)
            gR_[fs["perm"]] = gq
            RR = fs["IR"].T @ gR_
            for v in range(4):
                r[self.off[fs["eR"]] + v * NR * NR + face_nodes(NR, fs["fR"])] += RR[:, v]
        return r
