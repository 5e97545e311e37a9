"""2-D DG residual on meshes mixing TRI and QUAD elements of any order, in numpy -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline may import this module; the product path
never does.  It restates the reference's per-volume / per-face loops with a dense operator per element, for
the element types of the reference's mixed 2-D meshes:

* QUAD: Lagrange basis on the GLL nodes of order p (nodes r fastest), (p+1)^2 GL volume cubature.
* TRI: the reference's simplex (nodes.c:628-852): the equilateral triangle with vertices (-1, -1/sqrt 3),
  (1, -1/sqrt 3), (0, 2/sqrt 3); Lagrange basis on the WSH nodes of order p (interp_si WSH), WV volume
  cubature of degree 2p (nodes_operators.c:436-520; p_cub = 2p, compute_p_nodes :327-345); local face lf
  opposite vertex lf, running from vertex f_ve[lf][0] to f_ve[lf][1] = (1, 2), (0, 2), (0, 1) (element.c:454-463).
* Both: GL face cubature of the face order max(p_L, p_R) + 1 points (compute_face_reference_order,
  face_solver.c:123-131), each side's own test functions at those points (compute_face_rlhs_T.c:51-80), the
  face geometry n J_f from the lower-order side, left on a tie (get_dominant_geom_vol_side_index,
  face_solver.c:95-107); the right trace taken in the left ordering (ind_ord 1 reverses a LINE face).
* Volume term: R += sum_d tw1_d fr_d with the J-scaled cofactor metrics of the element's isoparametric map at
  its cubature (compute_volume_rlhs_T.c:145-177, geometry_T.c:546-590).
* Navier-Stokes (rhs_ns, viscous=(mu, Pr)): BR2 with the weak gradient, the face lifts at the face cubature
  and the penalty 1.01 * NFMAX, NFMAX = 4 for every 2-D element (definitions_elements.h:68,
  compute_face_rlhs_dg_T.c:275-296), and the central viscous face flux.

The TRI node and cubature tables are the reference's input files (input/nodes/tri/WSH*.txt, WV*.txt), read
from tests/golden/tri_nodes.json and expanded orbit by orbit in the reference's permutation order.  Pinned
by tests/test_mixed_oracle.py: the expanded nodes equal the golden nodes_si.data; the cubature integrates
monomials to its degree; on QUAD-only meshes the residual equals hp_oracle's; free-stream preservation and
discrete conservation on mixed meshes; convergence of M^-1 R to the exact -div F of a smooth state.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

import pyoracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(os.path.dirname(HERE), "tests", "golden", "tri_nodes.json")
QUAD, TRI = 0, 1
SQRT3 = math.sqrt(3.0)
TRI_V = np.array([[-1.0, -1.0 / SQRT3], [1.0, -1.0 / SQRT3], [0.0, 2.0 / SQRT3]])   # nodes.c:636-637
TRI_F_VE = ((1, 2), (0, 2), (0, 1))                                                 # element.c:461
_PERMS = ((0, 1, 2), (2, 0, 1), (1, 2, 0), (0, 2, 1), (1, 0, 2), (2, 1, 0))         # nodes.c:634
_tables = None


def tri_table(name):
    global _tables
    if _tables is None:
        _tables = json.load(open(GOLDEN))["tables"]
    return _tables[name]


def tri_nodes(kind, order):
    """(barycentric (n, 3), rst (n, 2), w (n,)) of input/nodes/tri/<kind><order>.txt, expanded as the reference
    expands its symmetry orbits: groups in file order, each orbit through the first Nperms permutations."""
    t = tri_table(f"{kind}{order}")
    symms = t["symms"]
    ig = 0
    while symms[ig][0] == 0:
        ig += 1
    count = 0
    bary, w = [], []
    for g, b in enumerate(t["bary"]):
        count += 1
        for j in range(symms[ig][1]):
            bary.append([b[_PERMS[j][k]] for k in range(3)])
            w.append(t["w"][g])
        if count == symms[ig][0]:
            count = 0
            ig += 1
    bary = np.array(bary)
    return bary, bary @ TRI_V, np.array(w)


_BARY = np.linalg.inv(np.vstack([TRI_V.T, np.ones(3)]))    # (r, s, 1) -> barycentric coordinates


def _pkd_values(rs, p):
    """Proriol-Koornwinder-Dubiner polynomials P_i(a) ((1-b)/2)^i P_j^(2i+1,0)(b), i + j <= p, at points rs (n, 2)
    (complex allowed), with x = 2 l1 - 1, y = 2 l2 - 1 the right-triangle coordinates of the barycentric l and
    (a, b) = (2 (1+x)/(1-y) - 1, y) the collapsed ones."""
    from scipy.special import eval_jacobi

    lam = np.einsum("ij,nj->ni", _BARY, np.column_stack([rs[:, 0], rs[:, 1], np.ones(len(rs))]))
    x, y = 2 * lam[:, 1] - 1, 2 * lam[:, 2] - 1
    a, b = 2 * (1 + x) / (1 - y) - 1, y
    cols = []
    for i in range(p + 1):
        for j in range(p + 1 - i):
            cols.append(eval_jacobi(i, 0.0, 0.0, a) * (0.5 * (1 - b)) ** i * eval_jacobi(j, 2 * i + 1.0, 0.0, b))
    return np.stack(cols, 1)


def _orthogonal_basis(rs, p):
    """The PKD basis and its r, s derivatives (complex step) at points rs (n, 2): an orthogonal basis of P_p,
    so the Vandermonde matrix of the nodes is well conditioned in float64."""
    rs = np.asarray(rs, float)
    h = 1e-30
    V = _pkd_values(rs.astype(complex), p).real
    Vr = _pkd_values(rs + np.array([1j * h, 0]), p).imag / h
    Vs = _pkd_values(rs + np.array([0, 1j * h]), p).imag / h
    return V, Vr, Vs


class TriBasis:
    """Lagrange basis of order p on the WSH nodes (through the orthogonal PKD basis; the operators are basis-free)."""

    def __init__(self, p):
        self.p = p
        _, self.nodes, _ = tri_nodes("WSH", p)
        self.Np = len(self.nodes)
        V, _, _ = _orthogonal_basis(self.nodes, p)
        self.Vinv = np.linalg.inv(V)

    def eval(self, rs):
        """(phi, dphi/dr, dphi/ds) at points rs (n, 2), each (n, Np)."""
        V, Vr, Vs = _orthogonal_basis(np.atleast_2d(rs), self.p)
        return V @ self.Vinv, Vr @ self.Vinv, Vs @ self.Vinv


class QuadBasis:
    """Lagrange basis of order p on the GLL nodes, node i + N j at (x_i, x_j)."""

    def __init__(self, p):
        self.p, self.N = p, p + 1
        self.x1, _ = O.gauss_lobatto(self.N)
        self.Np = self.N ** 2

    def eval(self, rs):
        rs = np.atleast_2d(rs)
        Ir, Dr = O.lagrange(self.x1, rs[:, 0])
        Is, Ds = O.lagrange(self.x1, rs[:, 1])
        phi = np.einsum("qi,qj->qji", Ir, Is).reshape(len(rs), -1)
        dr = np.einsum("qi,qj->qji", Dr, Is).reshape(len(rs), -1)
        ds = np.einsum("qi,qj->qji", Ir, Ds).reshape(len(rs), -1)
        return phi, dr, ds


_basis_cache = {}


def basis(kind, p):
    if (kind, p) not in _basis_cache:
        _basis_cache[(kind, p)] = TriBasis(p) if kind == TRI else QuadBasis(p)
    return _basis_cache[(kind, p)]


def volume_cubature(kind, p):
    if kind == TRI:
        _, rs, w = tri_nodes("WV", 2 * p)
        return rs, w
    xg, wg = O.gauss_legendre(p + 1)
    rs = np.stack([np.tile(xg, p + 1), np.repeat(xg, p + 1)], 1)
    return rs, np.kron(wg, wg)


def face_points(kind, lf, s):
    """Reference coordinates of face coordinate s in [-1, 1] on local face lf, and d(rs)/ds."""
    s = np.asarray(s, float)
    if kind == TRI:
        a, b = TRI_V[TRI_F_VE[lf][0]], TRI_V[TRI_F_VE[lf][1]]
        return np.outer(0.5 * (1 - s), a) + np.outer(0.5 * (1 + s), b), 0.5 * (b - a)
    fd, fix = lf // 2, (1.0 if lf % 2 else -1.0)
    pts = np.empty((len(s), 2))
    pts[:, fd] = fix
    pts[:, 1 - fd] = s
    t = np.zeros(2)
    t[1 - fd] = 1.0
    return pts, t


def face_sign(kind, lf):
    """+1 when the outward normal is the tangent turned clockwise, (t_y, -t_x), for a counter-clockwise element."""
    if kind == TRI:
        return -1.0 if lf == 1 else 1.0
    fd, side = lf // 2, lf % 2
    return (1.0 if side else -1.0) * (1.0 if fd == 0 else -1.0)


class MixedOracle:
    """Euler residual (and BR2 Navier-Stokes with viscous=(mu, Pr)) on a mixed mesh (dpgsolver_amd.mesh.MixedMesh:
    kind, p, xyz per element at its solution nodes, faces {eL, fL, eR, fR, ord, bc}, offsets(nv))."""

    def __init__(self, mesh, flux, bc_prim=None, viscous=None, walls=None):
        """walls = (nf_E, omega, rho_b, E_b): the Navier-Stokes no-slip wall data (dgx_config wall_*)."""
        self.mesh, self.flux, self.viscous = mesh, flux, viscous
        self.walls = (0.0, 0.0, 0.0, 0.0) if walls is None else tuple(float(w) for w in walls)
        self.nv = 4
        self.bc_prim = np.zeros(5) if bc_prim is None else np.resize(np.asarray(bc_prim, float), 5)
        self.off = mesh.offsets(self.nv)
        self._geo = [self._volume(e) for e in range(mesh.E)]

    # ---------------------------------------------------------------------------------------------- volumes
    def _volume(self, e):
        m = self.mesh
        kind, p = int(m.kind[e]), int(m.p[e])
        B = basis(kind, p)
        rs, W = volume_cubature(kind, p)
        I, Dr, Ds = B.eval(rs)
        x = m.xyz[e]
        xr, xs = x @ Dr.T, x @ Ds.T                                   # (2, Nc)
        J = xr[0] * xs[1] - xr[1] * xs[0]
        return dict(kind=kind, p=p, B=B, I=I, Dr=Dr, Ds=Ds, W=W, J=J, xr=xr, xs=xs, Np=B.Np)

    def element(self, u, e):
        return u[self.off[e]:self.off[e + 1]].reshape(self.nv, self._geo[e]["Np"])

    def mass(self, e):
        g = self._geo[e]
        return g["I"].T @ ((g["W"] * g["J"])[:, None] * g["I"])

    def apply_minv(self, r):
        out = np.array(r, float)
        for e in range(self.mesh.E):
            R = out[self.off[e]:self.off[e + 1]].reshape(self.nv, -1)
            out[self.off[e]:self.off[e + 1]] = np.linalg.solve(self.mass(e), R.T).T.ravel()
        return out

    def _volume_term(self, e, U, grad=None):
        g = self._geo[e]
        s = g["I"] @ U.T                                              # (Nc, nv)
        f = np.stack([O.euler_flux(2, s[q]).reshape(self.nv, 2) for q in range(len(s))])   # (Nc, nv, 2)
        if grad is not None:
            mu, Pr = self.viscous
            gq = np.einsum("qi,dvi->qdv", g["I"], grad)              # (Nc, d, nv)
            f = f + np.stack([O.viscous_flux(2, s[q], gq[q].reshape(-1), mu, Pr) for q in range(len(s))])
        xr, xs = g["xr"], g["xs"]
        fr0 = xs[1][:, None] * f[:, :, 0] - xs[0][:, None] * f[:, :, 1]
        fr1 = -xr[1][:, None] * f[:, :, 0] + xr[0][:, None] * f[:, :, 1]
        return (g["Dr"].T @ (g["W"][:, None] * fr0) + g["Ds"].T @ (g["W"][:, None] * fr1)).T   # (nv, Np)

    # ---------------------------------------------------------------------------------------------- faces
    def _side(self, e, lf, s):
        """Basis of element e at face coordinates s of its face lf, and n J of its own map there (outward)."""
        g = self._geo[e]
        pts, dpds = face_points(g["kind"], lf, s)
        phi, dr, ds = g["B"].eval(pts)
        t = self.mesh.xyz[e] @ (dr * dpds[0] + ds * dpds[1]).T          # (2, nq) dx/ds
        nJ = face_sign(g["kind"], lf) * np.stack([t[1], -t[0]])
        return phi, nJ

    def _face_setup(self, u, row):
        eL, fL, eR, fR, ord_, bc = (int(v) for v in row)
        m = self.mesh
        pL = int(m.p[eL])
        pR = pL if bc != 0 else int(m.p[eR])
        Nf = max(pL, pR) + 1
        sq, wq = O.gauss_legendre(Nf)
        IL, nJL = self._side(eL, fL, sq)
        UL = self.element(u, eL)
        sL = IL @ UL.T                                                  # (Nf, nv) left ordering
        out = dict(eL=eL, fL=fL, eR=eR, fR=fR, bc=bc, Nf=Nf, wq=wq, IL=IL, sL=sL)
        if bc != 0:
            out["nJ"] = nJL
            out["xy"] = IL @ m.xyz[eL].T                                # (Nf, 2) the face points (isoparametric)
            return out
        sR_pts = -sq if ord_ == 1 else sq                               # the right face coordinate of point q
        IR, nJR = self._side(eR, fR, sR_pts)
        out["IR"] = IR
        out["sR"] = IR @ self.element(u, eR).T
        out["nJ"] = nJL if pL <= pR else -nJR
        return out

    def rhs(self, u, part="all"):
        """Euler residual, no M^-1; part "volume" or "face" for one term alone (compute_volume_rlhs_dg /
        compute_face_rlhs_dg)."""
        u = np.asarray(u, float)
        r = np.zeros_like(u)
        if part != "face":
            for e in range(self.mesh.E):
                r[self.off[e]:self.off[e + 1]] += self._volume_term(e, self.element(u, e)).ravel()
        if part == "volume":
            return r
        for row in self.mesh.faces:
            fs = self._face_setup(u, row)
            nJ = fs["nJ"]
            Jf = np.sqrt((nJ ** 2).sum(0))
            n = nJ / Jf
            if fs["bc"] != 0:
                sR = np.stack([O.bc_state(2, fs["bc"], fs["sL"][q], n[:, q], self.bc_prim) for q in range(fs["Nf"])])
            else:
                sR = fs["sR"]
            nnf = np.stack([O.numflux(2, 1, self.flux, fs["sL"][q], sR[q], n[:, q]) for q in range(fs["Nf"])])
            gq = (fs["wq"] * Jf)[:, None] * nnf
            self._lift(r, fs, gq)
        return r

    def _lift(self, r, fs, gq):
        eL, eR = fs["eL"], fs["eR"]
        r[self.off[eL]:self.off[eL + 1]] -= (fs["IL"].T @ gq).T.ravel()
        if fs["bc"] == 0:
            r[self.off[eR]:self.off[eR + 1]] += (fs["IR"].T @ gq).T.ravel()

    # ---------------------------------------------------------------------------------------------- BR2
    def _ns_boundary_states(self, fs):
        """Navier-Stokes boundary states at a boundary face's points: Riemann (the Euler state); the no-slip
        adiabatic / diabatic walls keep rho and E and negate the momentum (boundary_navier_stokes_T.c:260-290);
        the rotating wall is 2 s_ex - s_l, s_ex = (rho_b, rho_b u_ex, E_b) with u_ex = omega r (-sin th, cos th)
        at the point (:200-240, :505-520)."""
        nf_E, omega, rho_b, E_b = self.walls
        n = fs["nJ"] / np.sqrt((fs["nJ"] ** 2).sum(0))
        out = np.empty_like(fs["sL"])
        for q in range(fs["Nf"]):
            sl = fs["sL"][q]
            if fs["bc"] == 1:
                out[q] = O.bc_state(2, 1, sl, n[:, q], self.bc_prim)
            elif fs["bc"] == 5:
                x, y = fs["xy"][q]
                r, th = np.hypot(x, y), np.arctan2(y, x)
                uex = np.array([-np.sin(th), np.cos(th)]) * omega * r
                out[q] = np.concatenate([[2 * rho_b - sl[0]], 2 * rho_b * uex - sl[1:3], [2 * E_b - sl[3]]])
            elif fs["bc"] in (3, 4):
                out[q] = np.array([sl[0], -sl[1], -sl[2], sl[3]])
            else:
                raise ValueError(f"boundary code {fs['bc']} has no Navier-Stokes condition")
        return out

    def ns_gradient(self, u):
        """(gfull, gv, gf): per element the weak gradient (d, nv, Np_e), its volume part, and per face the lifts
        of both sides (compute_grad_coef_dg_T.c:171-307, central s_num)."""
        m = self.mesh
        gv = []
        for e in range(m.E):
            g = self._geo[e]
            U = self.element(u, e)
            sr, ss = U @ g["Dr"].T, U @ g["Ds"].T                        # (nv, Nc)
            xr, xs = g["xr"], g["xs"]
            Jg = [xs[1] * sr - xr[1] * ss, -xs[0] * sr + xr[0] * ss]    # J ds/dx, J ds/dy
            Minv = np.linalg.inv(self.mass(e))
            gv.append(np.stack([Minv @ (g["I"].T @ (g["W"][:, None] * Jd.T)) for Jd in Jg]).transpose(0, 2, 1))
        gfull = [x.copy() for x in gv]
        gf = []
        for row in m.faces:
            fs = self._face_setup(u, row)
            if fs["bc"] != 0:   # s_num = (s_L + s_b) / 2 (compute_g_coef_related_boundary), left side only
                fs["sR"] = self._ns_boundary_states(fs)
            dl = 0.5 * (fs["sR"] - fs["sL"])
            pair = []
            for side in ((0, 1) if fs["bc"] == 0 else (0,)):
                e, I = (fs["eL"], fs["IL"]) if side == 0 else (fs["eR"], fs["IR"])
                Minv = np.linalg.inv(self.mass(e))
                lift = np.stack([(Minv @ (I.T @ ((fs["wq"] * fs["nJ"][dim])[:, None] * dl))).T for dim in range(2)])
                gfull[e] += lift
                pair.append(lift)
            gf.append(pair)
        return gfull, gv, gf

    def rhs_ns(self, u):
        mu, Pr = self.viscous
        eta = 1.01 * 4
        u = np.asarray(u, float)
        gfull, gv, gf = self.ns_gradient(u)
        r = np.zeros_like(u)
        for e in range(self.mesh.E):
            r[self.off[e]:self.off[e + 1]] += self._volume_term(e, self.element(u, e), gfull[e]).ravel()
        for k, row in enumerate(self.mesh.faces):
            fs = self._face_setup(u, row)
            gL = np.einsum("qi,dvi->qdv", fs["IL"], gv[fs["eL"]] + eta * gf[k][0])
            if fs["bc"] == 0:
                gR = np.einsum("qi,dvi->qdv", fs["IR"], gv[fs["eR"]] + eta * gf[k][1])
            else:               # the boundary state, and the interior gradient as the boundary one
                fs["sR"] = self._ns_boundary_states(fs)
                gR = gL
            Jf = np.sqrt((fs["nJ"] ** 2).sum(0))
            n = fs["nJ"] / Jf
            tot = np.empty((fs["Nf"], 4))
            for q in range(fs["Nf"]):
                inv = O.numflux(2, 1, self.flux, fs["sL"][q], fs["sR"][q], n[:, q])
                fvL = O.viscous_flux(2, fs["sL"][q], gL[q].reshape(-1), mu, Pr)
                fvR = O.viscous_flux(2, fs["sR"][q], gR[q].reshape(-1), mu, Pr)
                tot[q] = inv + 0.5 * ((fvL + fvR) @ n[:, q])
                if fs["bc"] in (3, 4):   # no-slip flux walls: the provided energy flux (numerical_flux_central_T.c)
                    tot[q][3] = self.walls[0]
            self._lift(r, fs, (fs["wq"] * Jf)[:, None] * tot)
        return r
