"""Structured tensor-product meshes, face tables and seeded states.

The reference reads gmsh meshes and matches periodic faces
(src/simulation/mesh/mesh_readers.c, mesh/mesh_periodic.c:102) and stores the
left/right face correspondence as ``neigh_info{volume, ind_lf, ind_ord}``
(src/simulation/computational_elements/face.h:32-70).  Here meshes are
generated in memory and flattened into the arrays the C-ABI takes:

* ``xyz``   (E, d, Np) float64 -- isoparametric geometry nodes at the GLL solution
  nodes, r-fastest ordering (element_operators_tp.c:90-131).
* ``faces`` (F, 6) int32 -- ``eL, fL, eR, fR, ord, bc`` with TP local faces
  ``2*dim + side`` (side 0 = -1, 1 = +1) and ``ord`` the reference's face
  orientation index ind_ord (``face_perm``).

Face node ordering: the face coordinates are the remaining volume dimensions in
increasing order, first fastest.  ``perm[ord][i]`` is the right-face index of the
node that coincides with left-face node ``i``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

BC_INTERIOR, BC_RIEMANN, BC_SLIPWALL = 0, 1, 2
# Navier-Stokes no-slip walls (definitions_bc.h: BC_NOSLIP_ADIABATIC 111, BC_NOSLIP_DIABATIC 112,
# BC_NOSLIP_ALL_ROTATING 113)
BC_NOSLIP_ADIABATIC, BC_NOSLIP_DIABATIC, BC_NOSLIP_ROTATING = 3, 4, 5
PDE_ADVECTION, PDE_EULER, PDE_NAVIER_STOKES = 0, 1, 2
FLUX_UPWIND, FLUX_LF, FLUX_ROE = 0, 1, 2
GAMMA, GM1 = 1.4, 0.4


# --------------------------------------------------------------------------- nodes
def gauss_lobatto(n: int) -> tuple[np.ndarray, np.ndarray]:
    """GLL nodes/weights on [-1, 1] (interior nodes: roots of P_{n-1}')."""
    N = n - 1
    P = np.polynomial.legendre.Legendre.basis(N)
    x = np.concatenate([[-1.0], np.sort(P.deriv().roots().real), [1.0]])
    # polish the interior roots with Newton on P_N'
    dP, d2P = P.deriv(), P.deriv(2)
    for _ in range(3):
        x[1:-1] -= dP(x[1:-1]) / d2P(x[1:-1])
    w = 2.0 / (N * (N + 1) * P(x) ** 2)
    x = 0.5 * (x - x[::-1])
    w = 0.5 * (w + w[::-1])
    return x, w


def gauss_legendre(n: int) -> tuple[np.ndarray, np.ndarray]:
    x, w = np.polynomial.legendre.leggauss(n)
    x = 0.5 * (x - x[::-1])
    w = 0.5 * (w + w[::-1])
    return x, w


def _tp_coords(N: int, d: int) -> np.ndarray:
    """(Np, d) integer multi-indices, dimension 0 fastest."""
    idx = np.arange(N ** d)
    return np.stack([(idx // N ** k) % N for k in range(d)], axis=1)


def face_node_volume_index(d: int, N: int, lf: int) -> np.ndarray:
    """Volume node index of each face-local node of local face lf."""
    fd, side = lf // 2, lf % 2
    nf = N ** (d - 1)
    out = np.empty(nf, dtype=np.int64)
    rem = [k for k in range(d) if k != fd]
    for i in range(nf):
        ab = [(i // N ** j) % N for j in range(d - 1)]
        idx = [0] * d
        for j, k in enumerate(rem):
            idx[k] = ab[j]
        idx[fd] = N - 1 if side else 0
        out[i] = sum(idx[k] * N ** k for k in range(d))
    return out


def face_perm(d: int, N: int, ord_: int) -> np.ndarray:
    """perm[i] = the right-face node that coincides with left-face node i, for the reference's orientation index
    ind_ord (the face node correspondence of src/element/nodes_correspondence.c: LINE faces 0 aligned, 1
    reversed; QUAD faces, node (a, b) with a fastest: cases 4-7 transposed, cases 1, 3, 6, 7 reverse a, cases
    2, 3, 5, 7 reverse b)."""
    nf = N ** (d - 1)
    perm = np.empty(nf, dtype=np.int64)
    if d == 1:
        perm[0] = 0
        return perm
    if d == 2:
        if ord_ not in (0, 1):
            raise ValueError(f"LINE face orientation {ord_} (0 or 1)")
        perm[:] = np.arange(N)[::-1] if ord_ == 1 else np.arange(N)
        return perm
    if not 0 <= ord_ < 8:
        raise ValueError(f"QUAD face orientation {ord_} (0..7)")
    for b in range(N):
        for a in range(N):
            a2 = N - 1 - a if ord_ in (1, 3, 6, 7) else a
            b2 = N - 1 - b if ord_ in (2, 3, 5, 7) else b
            perm[a + N * b] = (b2 + N * a2) if ord_ >= 4 else (a2 + N * b2)
    return perm


def n_orientations(d: int) -> int:
    return {1: 1, 2: 2, 3: 8}[d]


# --------------------------------------------------------------------------- rng
class SplitMix64:
    """splitmix64 (Steele, Lea & Flood 2014) -- the seeded generator for all synthetic inputs."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.state = seed & self.MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & self.MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self, n: int, lo: float = -1.0, hi: float = 1.0) -> np.ndarray:
        """n doubles in [lo, hi) from the top 53 bits; vectorised for large n."""
        if n <= 4096:
            v = np.array([(self.next_u64() >> 11) * (1.0 / (1 << 53)) for _ in range(n)])
            return lo + (hi - lo) * v
        # vectorised splitmix64 over a counter: state_k = s0 + k*gamma
        g = np.uint64(0x9E3779B97F4A7C15)
        with np.errstate(over="ignore"):
            ks = np.arange(1, n + 1, dtype=np.uint64)
            z = np.uint64(self.state) + ks * g
            self.state = int(z[-1])
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        v = (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
        return lo + (hi - lo) * v


SEED = 20261202


# --------------------------------------------------------------------------- mesh
@dataclass
class Mesh:
    d: int
    p: int
    xyz: np.ndarray            # (E, d, Np)
    faces: np.ndarray          # (F, 6) int32
    periods: tuple = ()
    dims: tuple = ()
    meta: dict = field(default_factory=dict)

    @property
    def N(self) -> int:
        return self.p + 1

    @property
    def E(self) -> int:
        return self.xyz.shape[0]

    @property
    def F(self) -> int:
        return self.faces.shape[0]

    @property
    def Np(self) -> int:
        return self.N ** self.d


def _match_faces(d, N, xyz, pairs, periods):
    """Given element pairs (eL, eR) that share a face, find (fL, fR, ord) geometrically."""
    nf = N ** (d - 1)
    fidx = [face_node_volume_index(d, N, lf) for lf in range(2 * d)]
    perms = [face_perm(d, N, o) for o in range(n_orientations(d))]
    per = np.array([p if p else 0.0 for p in periods] + [0.0] * (d - len(periods)))

    def wrapdiff(a, b):
        diff = a - b
        for k in range(d):
            if per[k] > 0:
                diff[..., k] -= per[k] * np.round(diff[..., k] / per[k])
        return diff

    out = []
    for eL, eR, k in pairs:
        cL = [xyz[eL][:, fidx[lf]].T for lf in range(2 * d)]  # (nf, d)
        cR = [xyz[eR][:, fidx[lf]].T for lf in range(2 * d)]
        found = None
        # the left face is the one of eL lying furthest along +k
        fL_hi = int(np.argmax([c[:, k].mean() for c in cL]))
        for fL in (fL_hi,):
            mL = cL[fL].mean(0)
            for fR in range(2 * d):
                if eL == eR and fL == fR:
                    continue
                if np.abs(wrapdiff(mL, cR[fR].mean(0))).max() > 1e-9:
                    continue
                for o, pm in enumerate(perms):
                    if np.abs(wrapdiff(cL[fL], cR[fR][pm])).max() < 1e-9:
                        found = (fL, fR, o)
                        break
                if found:
                    break
            if found:
                break
        if found is None:
            raise ValueError(f"faces of elements {eL},{eR} do not match")
        out.append(found)
    return out


def structured_tp_mesh(d: int, p: int, n: tuple, lo: tuple, hi: tuple, periodic: tuple,
                       curve=None, bc_low=None, bc_high=None, rotate=None) -> Mesh:
    """Structured QUAD/HEX mesh of prod(n) elements, i fastest.

    curve(x: (..., d)) -> (..., d) maps straight node coordinates to curved ones.
    rotate: optional array (E,) of orientation codes; element local numbering is
    permuted (2-D: 90-degree turns, 3-D: axis flips/permutations) to exercise
    general face orientations.
    """
    N = p + 1
    x1, _ = gauss_lobatto(N)
    ref = x1[_tp_coords(N, d)]                    # (Np, d) in [-1, 1]
    E = int(np.prod(n))
    eidx = _tp_coords_var(n)                      # (E, d)
    h = np.array([(hi[k] - lo[k]) / n[k] for k in range(d)])
    base = np.array(lo)[None, :] + eidx * h[None, :]
    pts = base[:, None, :] + 0.5 * (ref[None, :, :] + 1.0) * h[None, None, :]   # (E, Np, d)
    if curve is not None:
        pts = curve(pts)
    if rotate is not None:
        pts = _apply_rotation(d, N, pts, rotate)
    xyz = np.ascontiguousarray(pts.transpose(0, 2, 1))

    pairs, bfaces = [], []
    strides = [int(np.prod(n[:k])) for k in range(d)]
    all_e = np.arange(E)
    for k in range(d):
        ck = eidx[:, k]
        nxt = ck + 1
        inside = nxt < n[k]
        en = all_e + strides[k]
        if periodic[k]:
            en = np.where(inside, en, all_e - (n[k] - 1) * strides[k])
            sel = np.ones(E, dtype=bool)
        else:
            sel = inside
            bfaces += [(int(e), k, 1) for e in all_e[~inside]]
            bfaces += [(int(e), k, 0) for e in all_e[ck == 0]]
        pairs += list(zip(all_e[sel].tolist(), en[sel].tolist(), [k] * int(sel.sum())))
    pairs.sort(key=lambda t: (t[0], t[2]))
    periods = tuple((hi[k] - lo[k]) if periodic[k] else 0.0 for k in range(d))
    if rotate is None:
        # unrotated structured numbering: the +k face of eL meets the -k face of eR, same orientation
        rows = [(eL, 2 * k + 1, eR, 2 * k, 0, BC_INTERIOR) for (eL, eR, k) in pairs]
    else:
        matched = _match_faces(d, N, xyz, pairs, periods)
        rows = [(eL, fL, eR, fR, o, BC_INTERIOR) for (eL, eR, _), (fL, fR, o) in zip(pairs, matched)]
    if bfaces and rotate is not None:
        raise ValueError("rotated elements are only supported on fully periodic meshes")
    for e, k, side in bfaces:
        rows.append((e, 2 * k + side, -1, -1, 0, (bc_high if side else bc_low)[k]))
    faces = np.array(rows, dtype=np.int32).reshape(-1, 6)
    return Mesh(d, p, xyz, faces, periods, tuple(n))


def annulus_mesh(p: int, nr: int, nt: int, r_i: float = 0.5, r_o: float = 1.0, bc_inner: int = BC_NOSLIP_ROTATING,
                 bc_outer: int = BC_NOSLIP_ADIABATIC) -> Mesh:
    """QUAD mesh of the annulus r_i <= r <= r_o (the reference's Taylor-Couette geometry, n-cylinder_hollow with
    geometry_parameters.geo r_i = 0.5, r_o = 1), nr elements across and nt around, periodic in theta; the
    isoparametric GLL nodes lie on the exact circles.  Local xi runs outward, eta counter-clockwise; the faces on
    the inner / outer circle carry bc_inner / bc_outer."""
    def curve(pts):
        r = r_i + (r_o - r_i) * pts[..., 0]
        th = 2.0 * np.pi * pts[..., 1]
        return np.stack([r * np.cos(th), r * np.sin(th)], -1)
    return structured_tp_mesh(2, p, (nr, nt), (0.0, 0.0), (1.0, 1.0), (False, True), curve=curve,
                              bc_low=(bc_inner, BC_INTERIOR), bc_high=(bc_outer, BC_INTERIOR))


def taylor_couette_fields(x: np.ndarray, mu: float, Pr: float, omega: float = 1.0, r_i: float = 0.5, r_o: float = 1.0,
                          dTdr_o: float = 0.0, t_i: float = 1.0, p_i: float = 1.0):
    """The reference's Taylor-Couette test state (solution_taylor_couette_T.c, specific gas constant 1) at points
    x (..., 2): V_theta = c (1/r - r/r_o^2) with c = omega / (1/r_i^2 - 1/r_o^2), the conduction-dissipation
    temperature profile with dT/dr = dTdr_o at r_o, uniform pressure p_i, rho = p / T.  Returns (conserved state,
    rate), both (..., 4): rate is -div F of that state, rho V^2 / r e_r in the momentum equations and 0 otherwise
    (the uniform pressure leaves the centripetal term unbalanced; the theta-momentum and energy equations hold
    exactly)."""
    X, Y = x[..., 0], x[..., 1]
    kappa = mu * GAMMA / GM1 / Pr
    c = omega / (1.0 / r_i ** 2 - 1.0 / r_o ** 2)
    r, th = np.sqrt(X * X + Y * Y), np.arctan2(Y, X)
    Vt = c * (1.0 / r - r / r_o ** 2)
    t = (t_i - 2.0 * c * c / r_o ** 2 * mu / kappa * np.log(r / r_i) - c * c * mu / kappa * (1.0 / r ** 2 - 1.0 / r_i ** 2)
         + dTdr_o * r_o * np.log(r / r_i))
    rho = p_i / t
    prim = np.stack([rho, -np.sin(th) * Vt, np.cos(th) * Vt, p_i + 0 * X], -1)
    rate = np.stack([0 * X, rho * Vt ** 2 / r * np.cos(th), rho * Vt ** 2 / r * np.sin(th), 0 * X], -1)
    return prim_to_cons(prim, 2), rate


def taylor_couette_state(mesh: Mesh, mu: float, Pr: float, **kw):
    """taylor_couette_fields at the nodes of a QUAD mesh, packed: (state, rate)."""
    cons, rate = taylor_couette_fields(nodes_xyz(mesh), mu, Pr, **kw)
    return pack_state(cons), pack_state(rate)


def taylor_couette_walls(mu: float, Pr: float, omega: float = 1.0, r_i: float = 0.5, dTdr_o: float = 0.0,
                         t_i: float = 1.0, p_i: float = 1.0) -> tuple:
    """(nf_E, omega, rho_b, E_b) of the Taylor-Couette walls: the outer wall's normal viscous energy flux
    -kappa dT/dn (get_normal_flux_Energy, kappa = mu Cp / Pr), the inner wall's rotation, density p_i / t_i and
    total energy p_i / (gamma - 1) + rho (omega r_i)^2 / 2 (solution.data: rho_b 1, E_b 2.625)."""
    kappa = mu * GAMMA / GM1 / Pr
    rho_b = p_i / t_i
    return (-kappa * dTdr_o, omega, rho_b, p_i / GM1 + 0.5 * rho_b * (omega * r_i) ** 2)


def _tp_coords_var(n: tuple) -> np.ndarray:
    E = int(np.prod(n))
    idx = np.arange(E)
    out, stride = [], 1
    for k in range(len(n)):
        out.append((idx // stride) % n[k])
        stride *= n[k]
    return np.stack(out, axis=1)


def _apply_rotation(d, N, pts, rotate):
    """Renumber element-local nodes: pts[e, new_local] = pts[e, old_local(new_local)]."""
    coords = _tp_coords(N, d)
    out = pts.copy()
    for e in range(pts.shape[0]):
        r = int(rotate[e])
        if r == 0:
            continue
        c = coords.copy()
        if d == 2:
            for _ in range(r % 4):                 # 90-degree turn of the local frame
                c = np.stack([N - 1 - c[:, 1], c[:, 0]], axis=1)
        else:
            # proper rotations only (a mirrored local frame would have det J < 0): each bit is a
            # half turn about a face diagonal or an axis, so every face orientation code occurs
            if r & 1:                              # swap axes 0,1 and reverse axis 2
                c = c[:, [1, 0, 2]]
                c[:, 2] = N - 1 - c[:, 2]
            if r & 2:                              # swap axes 1,2 and reverse axis 0
                c = c[:, [0, 2, 1]]
                c[:, 0] = N - 1 - c[:, 0]
            if r & 4:                              # reverse axes 0 and 1
                c[:, 0] = N - 1 - c[:, 0]
                c[:, 1] = N - 1 - c[:, 1]
        src = sum(c[:, k] * N ** k for k in range(d))
        out[e] = pts[e][src]
    return out


# --------------------------------------------------------------------------- states
def prim_to_cons(prim: np.ndarray, d: int) -> np.ndarray:
    """prim[..., (rho, u.., p)] -> conserved."""
    rho = prim[..., 0]
    vel = prim[..., 1:1 + d]
    p = prim[..., d + 1]
    out = np.empty_like(prim)
    out[..., 0] = rho
    out[..., 1:1 + d] = rho[..., None] * vel
    out[..., d + 1] = p / GM1 + 0.5 * rho * (vel ** 2).sum(-1)
    return out


def nodes_xyz(mesh: Mesh) -> np.ndarray:
    """(E, Np, d) physical coordinates of the solution nodes (isoparametric)."""
    return mesh.xyz.transpose(0, 2, 1)


def pack_state(cons_nodes: np.ndarray) -> np.ndarray:
    """(E, Np, nv) -> flat reference layout per element (Np, nv) column-major."""
    return np.ascontiguousarray(cons_nodes.transpose(0, 2, 1)).ravel()


# --------------------------------------------------------------------------- configs
def config1_mesh(n: int = 16, p: int = 2) -> Mesh:
    """2-D advection, periodic [-1,1]^2 QUAD."""
    return structured_tp_mesh(2, p, (n, n), (-1.0, -1.0), (1.0, 1.0), (True, True))


def config1_state(mesh: Mesh) -> np.ndarray:
    x = nodes_xyz(mesh)
    u = np.sin(np.pi * x[..., 0]) * np.sin(np.pi * x[..., 1])
    return pack_state(u[..., None])


def config2_mesh(n: int = 128, p: int = 3) -> Mesh:
    """2-D Euler, periodic [0,10]^2 QUAD."""
    return structured_tp_mesh(2, p, (n, n), (0.0, 0.0), (10.0, 10.0), (True, True))


def isentropic_vortex(x: np.ndarray, beta=5.0, xc=(5.0, 5.0), mach=0.5) -> np.ndarray:
    """Isentropic vortex (strength beta, unit radius) on a Mach `mach` free stream along x; primitive."""
    dx = x[..., 0] - xc[0]
    dy = x[..., 1] - xc[1]
    r2 = dx * dx + dy * dy
    T = 1.0 - GM1 * beta ** 2 / (8.0 * GAMMA * math.pi ** 2) * np.exp(1.0 - r2)
    rho = T ** (1.0 / GM1)
    p = rho ** GAMMA
    uinf = mach * math.sqrt(GAMMA)             # rho_inf = p_inf = 1 -> c_inf = sqrt(gamma)
    u = uinf - beta / (2.0 * math.pi) * np.exp(0.5 * (1.0 - r2)) * dy
    v = beta / (2.0 * math.pi) * np.exp(0.5 * (1.0 - r2)) * dx
    return np.stack([rho, u, v, p], axis=-1)


def config2_state(mesh: Mesh, noise: float = 1e-3, seed: int = SEED) -> np.ndarray:
    x = nodes_xyz(mesh)
    cons = prim_to_cons(isentropic_vortex(x), 2)
    rng = SplitMix64(seed + 2)
    cons = cons + noise * rng.uniform(cons.size).reshape(cons.shape)
    return pack_state(cons)


def curve3d(amp: float = 0.05):
    def f(pts):
        x, y, z = pts[..., 0], pts[..., 1], pts[..., 2]
        s = 2.0 * np.pi
        out = pts.copy()
        out[..., 0] = x + amp * np.sin(s * y) * np.sin(s * z)
        out[..., 1] = y + amp * np.sin(s * z) * np.sin(s * x)
        out[..., 2] = z + amp * np.sin(s * x) * np.sin(s * y)
        return out
    return f


def config3_mesh(n: int = 64, p: int = 4, amp: float = 0.05) -> Mesh:
    """3-D Euler, periodic [0,1]^3 HEX, curved by x += amp sin(2 pi y) sin(2 pi z) (cyclic)."""
    return structured_tp_mesh(3, p, (n, n, n), (0.0,) * 3, (1.0,) * 3, (True,) * 3, curve=curve3d(amp))


def smooth_state_3d(x: np.ndarray, seed: int = SEED) -> np.ndarray:
    """rho = 1+0.1 s1, (u,v,w) = 0.28 (s2,s3,s4), p = 1 + 0.1 s5 with s_k = prod_c sin(2 pi (c + phi_kc))."""
    rng = SplitMix64(seed + 3)
    phi = rng.uniform(15, 0.0, 1.0).reshape(5, 3)
    d = x.shape[-1]
    # sin(2 pi (x + phi)) = sin(2 pi x) cos(2 pi phi) + cos(2 pi x) sin(2 pi phi): 2d transcendentals in total
    sx = [np.sin(2.0 * np.pi * x[..., c]) for c in range(d)]
    cx = [np.cos(2.0 * np.pi * x[..., c]) for c in range(d)]
    s = []
    for k in range(5):
        v = None
        for c in range(d):
            t = sx[c] * math.cos(2.0 * math.pi * phi[k, c]) + cx[c] * math.sin(2.0 * math.pi * phi[k, c])
            v = t if v is None else v * t
        s.append(v)
    prim = np.stack([1.0 + 0.1 * s[0], 0.28 * s[1], 0.28 * s[2], 0.28 * s[3], 1.0 + 0.1 * s[4]], axis=-1)
    if d == 2:
        prim = prim[..., [0, 1, 2, 4]]
    return prim


def config4_mesh(n: int = 448, p: int = 3) -> Mesh:
    """2-D Navier-Stokes workload on QUADs of one order p: periodic [0,1]^2, n x n elements (the per-element
    orders are config4_hp_mesh; the configuration's TRI half is not built)."""
    return structured_tp_mesh(2, p, (n, n), (0.0, 0.0), (1.0, 1.0), (True, True))


def config4_state(mesh: Mesh, seed: int = SEED) -> np.ndarray:
    """The configuration-3 state formula in 2-D, sub-seed 4."""
    return config3_state(mesh, seed + 1)


def config3_state(mesh: Mesh, seed: int = SEED) -> np.ndarray:
    """smooth_state_3d in conserved variables, built directly in the packed [E][nv][Np] layout."""
    d = mesh.d
    rng = SplitMix64(seed + 3)
    phi = rng.uniform(15, 0.0, 1.0).reshape(5, 3)
    xc = [np.ascontiguousarray(mesh.xyz[:, c, :]) for c in range(d)]
    sx = [np.sin(2.0 * np.pi * a) for a in xc]
    cx = [np.cos(2.0 * np.pi * a) for a in xc]
    s = []
    for k in range(5):
        v = None
        for c in range(d):
            t = sx[c] * math.cos(2.0 * math.pi * phi[k, c]) + cx[c] * math.sin(2.0 * math.pi * phi[k, c])
            v = t if v is None else v * t
        s.append(v)
    del sx, cx, xc
    rho = 1.0 + 0.1 * s[0]
    vel = [0.28 * s[1 + c] for c in range(d)]
    p = 1.0 + 0.1 * s[4]
    out = np.empty((mesh.E, d + 2, mesh.Np))
    out[:, 0] = rho
    ke = np.zeros_like(rho)
    for c in range(d):
        out[:, 1 + c] = rho * vel[c]
        ke += vel[c] * vel[c]
    out[:, d + 1] = p / GM1 + 0.5 * rho * ke
    return out.ravel()


def channel_mesh(nx: int = 64, ny: int = 64, p: int = 3) -> Mesh:
    """[-1.5,1.5] x [0,0.8] channel, bottom y = 0.0625 exp(-25 x^2); Riemann in/outflow, slip walls."""
    H = 0.8

    def curve(pts):
        out = pts.copy()
        x, eta = pts[..., 0], pts[..., 1] / H
        b = 0.0625 * np.exp(-25.0 * x * x)
        out[..., 1] = b + eta * (H - b)
        return out

    return structured_tp_mesh(2, p, (nx, ny), (-1.5, 0.0), (1.5, H), (False, False), curve=curve,
                              bc_low=(BC_RIEMANN, BC_SLIPWALL), bc_high=(BC_RIEMANN, BC_SLIPWALL))


def freestream_prim(d: int, mach: float = 0.5) -> np.ndarray:
    prim = np.zeros(d + 2)
    prim[0] = 1.0
    prim[1] = mach * math.sqrt(GAMMA)
    prim[d + 1] = 1.0
    return prim


def uniform_state(mesh: Mesh, prim: np.ndarray) -> np.ndarray:
    cons = prim_to_cons(prim[None, None, :], mesh.d)
    return pack_state(np.broadcast_to(cons, (mesh.E, mesh.Np, mesh.d + 2)).copy())


# --------------------------------------------------------------------------- variable order (2-D QUAD)
@dataclass
class HpMesh:
    """2-D QUAD mesh with one order per element (face_solver.c:123-131 gives a face the larger order of its
    two volumes).  Elements are sorted by order (non-decreasing), so the elements of one order form one
    contiguous group of the state vector, each in the uniform layout [var][node] of its own Np.
    * ``p``     (E,) int32 -- element orders, non-decreasing.
    * ``xyz``   list of (2, Np_e) float64 -- GLL geometry nodes per element, r-fastest.
    * ``faces`` (F, 6) int32 -- the uniform mesh's face table (eL, fL, eR, fR, ord, bc).
    * ``grid``  (E,) int -- structured-grid index of each element (before sorting by order)."""
    p: np.ndarray
    xyz: list
    faces: np.ndarray
    periods: tuple = ()
    dims: tuple = ()
    grid: np.ndarray = None
    d: int = 2

    @property
    def E(self) -> int:
        return len(self.p)

    @property
    def F(self) -> int:
        return self.faces.shape[0]

    def Np(self, e: int) -> int:
        return (int(self.p[e]) + 1) ** 2

    def offsets(self, nv: int) -> np.ndarray:
        """(E+1,) start of each element's block in the flat state vector."""
        sizes = (self.p.astype(np.int64) + 1) ** 2 * nv
        return np.concatenate([[0], np.cumsum(sizes)])

    def flat_xyz(self) -> np.ndarray:
        return np.ascontiguousarray(np.concatenate([x.ravel() for x in self.xyz]))


def structured_hp_mesh(p_elem, n: tuple, lo: tuple, hi: tuple, periodic: tuple, curve=None, rotate=None,
                       bc_low=None, bc_high=None) -> HpMesh:
    """Structured 2-D QUAD grid (i fastest) with per-element orders p_elem (grid order), reordered so that the
    orders are non-decreasing (stable).  Each element carries its own isoparametric GLL geometry nodes; a face
    is matched by its two end nodes, which every order shares (ord 0 or 1: same or reversed direction)."""
    d = 2
    p_grid = np.asarray(p_elem, dtype=np.int32)
    E = int(np.prod(n))
    assert p_grid.shape == (E,)
    eidx = _tp_coords_var(n)
    h = np.array([(hi[k] - lo[k]) / n[k] for k in range(d)])
    base = np.array(lo)[None, :] + eidx * h[None, :]
    xyz_grid = [None] * E
    for p in np.unique(p_grid):
        N = int(p) + 1
        x1, _ = gauss_lobatto(N)
        ref = x1[_tp_coords(N, d)]
        sel = np.nonzero(p_grid == p)[0]
        pts = base[sel][:, None, :] + 0.5 * (ref[None] + 1.0) * h[None, None, :]    # (E_p, Np, d)
        if curve is not None:
            pts = curve(pts)
        if rotate is not None:
            pts = _apply_rotation(d, N, pts, np.asarray(rotate)[sel])
        for k, e in enumerate(sel):
            xyz_grid[e] = np.ascontiguousarray(pts[k].T)
    pairs, bfaces = [], []
    strides = [int(np.prod(n[:k])) for k in range(d)]
    all_e = np.arange(E)
    for k in range(d):
        ck = eidx[:, k]
        inside = ck + 1 < n[k]
        en = all_e + strides[k]
        if periodic[k]:
            en = np.where(inside, en, all_e - (n[k] - 1) * strides[k])
            sel = np.ones(E, dtype=bool)
        else:
            sel = inside
            bfaces += [(int(e), k, 1) for e in all_e[~inside]]
            bfaces += [(int(e), k, 0) for e in all_e[ck == 0]]
        pairs += list(zip(all_e[sel].tolist(), en[sel].tolist(), [k] * int(sel.sum())))
    pairs.sort(key=lambda t: (t[0], t[2]))
    per = np.array([(hi[k] - lo[k]) if periodic[k] else 0.0 for k in range(d)])

    def wrap(a):
        a = a.copy()
        for k in range(d):
            if per[k] > 0:
                a[..., k] -= per[k] * np.round(a[..., k] / per[k])
        return a

    def ends(e, lf):
        N = int(p_grid[e]) + 1
        idx = face_node_volume_index(d, N, lf)
        return xyz_grid[e][:, idx[0]], xyz_grid[e][:, idx[-1]]

    def center(e, lf):
        a, b = ends(e, lf)
        return 0.5 * (a + b)

    rows = []
    if rotate is None:
        # unrotated structured numbering: the +k face of eL meets the -k face of eR in the same direction
        rows = [(eL, 2 * k + 1, eR, 2 * k, 0, BC_INTERIOR) for (eL, eR, k) in pairs]
        pairs = []
    for eL, eR, k in pairs:
        fL = int(np.argmax([center(eL, lf)[k] for lf in range(4)]))
        aL, bL = ends(eL, fL)
        found = None
        for fR in range(4):
            if eL == eR and fR == fL:
                continue
            aR, bR = ends(eR, fR)
            if np.abs(wrap(aL - aR)).max() < 1e-9 and np.abs(wrap(bL - bR)).max() < 1e-9:
                found = (fR, 0)
            elif np.abs(wrap(aL - bR)).max() < 1e-9 and np.abs(wrap(bL - aR)).max() < 1e-9:
                found = (fR, 1)
            if found:
                break
        if found is None:
            raise ValueError(f"faces of elements {eL},{eR} do not match")
        rows.append((eL, fL, eR, found[0], found[1], BC_INTERIOR))
    if bfaces and rotate is not None:
        raise ValueError("rotated elements are only supported on fully periodic meshes")
    for e, k, side in bfaces:
        rows.append((e, 2 * k + side, -1, -1, 0, (bc_high if side else bc_low)[k]))
    # sort the elements by order and renumber the faces
    order = np.argsort(p_grid, kind="stable")
    new_of = np.empty(E, dtype=np.int64)
    new_of[order] = np.arange(E)
    faces = np.array(rows, dtype=np.int64).reshape(-1, 6)
    faces[:, 0] = new_of[faces[:, 0]]
    inter = faces[:, 2] >= 0
    faces[inter, 2] = new_of[faces[inter, 2]]
    return HpMesh(p=p_grid[order].copy(), xyz=[xyz_grid[e] for e in order], faces=faces.astype(np.int32),
                  periods=tuple(per.tolist()), dims=tuple(n), grid=order)


def hp_nodes_xyz(mesh: HpMesh) -> list:
    """Per element (Np_e, 2) node coordinates."""
    return [x.T for x in mesh.xyz]


def hp_state(mesh: HpMesh, cons_fn) -> np.ndarray:
    """Flat state: cons_fn(x (n, 2)) -> (n, nv) conserved values (pointwise), packed per element [var][node]."""
    out = []
    for p in np.unique(mesh.p):
        sel = np.nonzero(mesh.p == p)[0]                    # one contiguous group (orders are sorted)
        x = np.stack([mesh.xyz[e].T for e in sel])          # (E_p, Np, 2)
        c = cons_fn(x.reshape(-1, 2)).reshape(len(sel), x.shape[1], -1)
        out.append(c.transpose(0, 2, 1).ravel())
    return np.ascontiguousarray(np.concatenate(out))


def config4_hp_orders(n: tuple, seed: int = SEED) -> np.ndarray:
    """Per-element orders p ~ U{1..5} of the variable-order Navier-Stokes workload (seeded, grid order)."""
    r = SplitMix64(seed + 4).uniform(int(np.prod(n)), 0.0, 1.0)
    return (1 + np.minimum((r * 5).astype(np.int32), 4)).astype(np.int32)


def config4_hp_mesh(nx: int = 500, ny: int = 300, seed: int = SEED) -> HpMesh:
    """Variable-order 2-D Navier-Stokes workload: the configuration's 500 x 300 base grid on periodic [0,1]^2,
    all QUAD (its seeded TRI split is not built), orders p ~ U{1..5} per element (seeded)."""
    return structured_hp_mesh(config4_hp_orders((nx, ny), seed), (nx, ny), (0.0, 0.0), (1.0, 1.0), (True, True))


def config4_hp_state(mesh: HpMesh, seed: int = SEED) -> np.ndarray:
    """The configuration-3 state formula in 2-D (as config4_state) at every element's own nodes."""
    return hp_state(mesh, lambda x: prim_to_cons(smooth_state_3d(x, seed + 1), 2))


# --------------------------------------------------------------------------- mixed TRI / QUAD (2-D)
QUAD_T, TRI_T = 0, 1
_TRI_F_VE = ((1, 2), (0, 2), (0, 1))        # TRI local face lf runs from vertex f_ve[lf][0] to f_ve[lf][1]
_QUAD_F_VE = ((0, 2), (1, 3), (0, 1), (2, 3))   # QUAD vertices (0,0) (1,0) (0,1) (1,1), face coordinate increasing


def tri_reference_nodes(p: int) -> np.ndarray:
    """(Np, 3) barycentric coordinates of the order-p TRI solution nodes (WSH, the reference's order), from the
    library (dgx_reference_nodes): the state at the ABI is nodal at these points."""
    from . import _lib

    rst, _ = _lib.reference_nodes(TRI_T, p)
    sq3 = math.sqrt(3.0)
    V = np.array([[-1.0, 1.0, 0.0], [-1.0 / sq3, -1.0 / sq3, 2.0 / sq3], [1.0, 1.0, 1.0]])
    return np.linalg.solve(V, np.vstack([rst.T, np.ones(len(rst))])).T


@dataclass
class MixedMesh:
    """2-D mesh of QUAD (type 0, GLL nodes) and TRI (type 1, WSH nodes) elements, sorted by (type, p) so that
    every (type, p) group is one contiguous range of the state.
    * ``kind``  (E,) int32, ``p`` (E,) int32;
    * ``xyz``   list of (2, Np_e) float64 geometry nodes at each element's solution nodes;
    * ``faces`` (F, 6) int32 (eL, fL, eR, fR, ord, bc), ord the reference's ind_ord (0 aligned, 1 reversed)."""
    kind: np.ndarray
    p: np.ndarray
    xyz: list
    faces: np.ndarray
    d: int = 2

    @property
    def E(self) -> int:
        return len(self.p)

    @property
    def F(self) -> int:
        return self.faces.shape[0]

    def Np(self, e: int) -> int:
        p = int(self.p[e])
        return (p + 1) * (p + 2) // 2 if self.kind[e] == TRI_T else (p + 1) ** 2

    def offsets(self, nv: int) -> np.ndarray:
        sizes = np.array([self.Np(e) * nv for e in range(self.E)], dtype=np.int64)
        return np.concatenate([[0], np.cumsum(sizes)])

    def flat_xyz(self) -> np.ndarray:
        return np.ascontiguousarray(np.concatenate([x.ravel() for x in self.xyz]))


def mixed_mesh_from_gmsh(path: str, p: int):
    """A gmsh 2.2 TRI / QUAD / mixed mesh (dgx_mesh_read_gmsh_mixed) as a MixedMesh of order p sorted by type
    (QUAD first, as dgx_upload_mesh_mixed wants), and the file index of each sorted element."""
    from . import _lib

    m = _lib.read_gmsh_mixed(path, p)
    kind = np.asarray(m["type"], np.int32)
    order = np.argsort(kind, kind="stable")              # sorted position -> file index
    where = np.empty_like(order)
    where[order] = np.arange(len(order))                 # file index -> sorted position
    faces = np.array(m["faces"], np.int32)
    faces[:, 0] = where[faces[:, 0]]
    inter = faces[:, 2] >= 0
    faces[inter, 2] = where[faces[inter, 2]]
    mesh = MixedMesh(kind=kind[order], p=np.full(len(order), p, np.int32), xyz=[m["xyz"][e] for e in order],
                     faces=faces)
    return mesh, order


def structured_mixed_mesh(n: tuple, split, p_elem, lo=(0.0, 0.0), hi=(1.0, 1.0), periodic=(True, True),
                          curve=None, bc_low=None, bc_high=None) -> MixedMesh:
    """Structured nx x ny grid (i fastest) whose cells with split[cell] are cut along the (i, j)-(i+1, j+1)
    diagonal into the counter-clockwise TRIs (c00, c10, c11) and (c00, c11, c01).  p_elem gives the order of each
    element in creation order (a QUAD cell: one; a split cell: two, A then B).  Faces are matched on the integer
    grid vertices (periodic directions wrap), the earlier-created element is the left side, ord 1 when the two
    sides run the other way; unmatched faces take bc_low / bc_high of their boundary (BC_* per direction).
    Geometry: each element's solution nodes mapped by the straight cell / triangle map, then by curve."""
    nx, ny = n
    split = np.asarray(split, bool).reshape(-1)
    h = ((hi[0] - lo[0]) / nx, (hi[1] - lo[1]) / ny)
    elems = []   # (kind, vertices as integer grid points)
    for j in range(ny):
        for i in range(nx):
            c00, c10, c01, c11 = (i, j), (i + 1, j), (i, j + 1), (i + 1, j + 1)
            if split[i + nx * j]:
                elems.append((TRI_T, (c00, c10, c11)))
                elems.append((TRI_T, (c00, c11, c01)))
            else:
                elems.append((QUAD_T, (c00, c10, c01, c11)))
    p_elem = np.asarray(p_elem, dtype=np.int32)
    if len(p_elem) != len(elems):
        raise ValueError(f"{len(elems)} elements, {len(p_elem)} orders")

    def wrap(v):
        x, y = v
        if periodic[0]:
            x %= nx
        if periodic[1]:
            y %= ny
        return (x, y)

    def phys(pts):
        out = np.empty_like(pts)
        out[..., 0] = lo[0] + pts[..., 0] * h[0]
        out[..., 1] = lo[1] + pts[..., 1] * h[1]
        return curve(out) if curve is not None else out

    # vertex weights of every element's solution nodes, per (type, p): barycentric (TRI) or bilinear (QUAD)
    weights = {}
    for kind, p in set(zip((k for k, _ in elems), p_elem.tolist())):
        if kind == TRI_T:
            weights[(kind, p)] = tri_reference_nodes(p)
        else:
            x1, _ = gauss_lobatto(p + 1)
            t = 0.5 * (x1 + 1.0)
            a, b = np.meshgrid(t, t)                      # a: r (fastest), b: s
            a, b = a.ravel(), b.ravel()
            weights[(kind, p)] = np.stack([(1 - a) * (1 - b), a * (1 - b), (1 - a) * b, a * b], 1)
    xyz = [None] * len(elems)
    groups = {}
    for e, ((kind, vs), p) in enumerate(zip(elems, p_elem.tolist())):
        groups.setdefault((kind, p), []).append(e)
    for key, ids in groups.items():
        V = np.array([elems[e][1] for e in ids], float)          # (n, nvert, 2)
        pts = phys(np.einsum("qv,evd->eqd", weights[key], V))   # (n, Np, 2)
        for k, e in enumerate(ids):
            xyz[e] = np.ascontiguousarray(pts[k].T)
    # faces: match on wrapped integer endpoints
    open_faces, rows = {}, []
    for e, (kind, vs) in enumerate(elems):
        fve = _TRI_F_VE if kind == TRI_T else _QUAD_F_VE
        for lf, (a, b) in enumerate(fve):
            A, B = wrap(vs[a]), wrap(vs[b])
            key = frozenset((A, B))
            if key in open_faces:
                e0, lf0, A0, B0 = open_faces.pop(key)
                rows.append([e0, lf0, e, lf, 0 if (A0 == A and B0 == B) else 1, BC_INTERIOR])
            else:
                open_faces[key] = (e, lf, A, B)
    for (e, lf, A, B) in open_faces.values():
        kind, vs = elems[e]
        fve = _TRI_F_VE if kind == TRI_T else _QUAD_F_VE
        va, vb = vs[fve[lf][0]], vs[fve[lf][1]]
        if va[0] == vb[0]:
            k, side = 0, (va[0] == nx)
        else:
            k, side = 1, (va[1] == ny)
        bc = (bc_high if side else bc_low)
        if bc is None:
            raise ValueError(f"unmatched face {e}/{lf} on a boundary without a condition")
        rows.append([e, lf, -1, -1, 0, bc[k]])
    # sort the elements by (type, p), stable, and renumber the faces
    kinds = np.array([k for k, _ in elems], dtype=np.int32)
    order = np.lexsort((p_elem, kinds))
    new_of = np.empty(len(elems), dtype=np.int64)
    new_of[order] = np.arange(len(elems))
    faces = np.array(rows, dtype=np.int64).reshape(-1, 6)
    faces[:, 0] = new_of[faces[:, 0]]
    inter = faces[:, 2] >= 0
    faces[inter, 2] = new_of[faces[inter, 2]]
    return MixedMesh(kind=kinds[order].copy(), p=p_elem[order].copy(), xyz=[xyz[e] for e in order],
                     faces=faces.astype(np.int32))


def annulus_mixed_mesh(n: tuple, split, p_elem, r_i: float = 0.5, r_o: float = 1.0,
                       bc_inner: int = BC_NOSLIP_ROTATING, bc_outer: int = BC_NOSLIP_ADIABATIC) -> MixedMesh:
    """Mixed TRI/QUAD mesh of the annulus r_i <= r <= r_o (the reference's Taylor-Couette geometry): the
    structured_mixed_mesh of n = (nr, ntheta) cells, periodic around, mapped onto the annulus (the solution nodes
    on the exact circles); the inner / outer circle's faces carry bc_inner / bc_outer."""
    def curve(pts):
        r = r_i + (r_o - r_i) * pts[..., 0]
        th = 2.0 * np.pi * pts[..., 1]
        return np.stack([r * np.cos(th), r * np.sin(th)], -1)
    return structured_mixed_mesh(n, split, p_elem, lo=(0.0, 0.0), hi=(1.0, 1.0), periodic=(False, True), curve=curve,
                                 bc_low=(bc_inner, BC_INTERIOR), bc_high=(bc_outer, BC_INTERIOR))


def taylor_couette_mixed_state(mesh: MixedMesh, mu: float, Pr: float, **kw):
    """taylor_couette_fields at the nodes of a mixed mesh: (state, rate) in the mixed_state layout."""
    return (mixed_state(mesh, lambda x: taylor_couette_fields(x, mu, Pr, **kw)[0]),
            mixed_state(mesh, lambda x: taylor_couette_fields(x, mu, Pr, **kw)[1]))


def mixed_state(mesh: MixedMesh, cons_fn) -> np.ndarray:
    """Flat state: cons_fn(x (n, 2)) -> (n, nv) conserved values at every element's own nodes, [var][node]
    (evaluated per (type, p) group: the elements are sorted by it)."""
    out = []
    key = mesh.kind.astype(np.int64) * 8 + mesh.p
    starts = np.concatenate([[0], np.nonzero(np.diff(key))[0] + 1, [mesh.E]])
    for a, b in zip(starts[:-1], starts[1:]):
        x = np.stack([mesh.xyz[e].T for e in range(a, b)])          # (n, Np, 2)
        c = cons_fn(x.reshape(-1, 2)).reshape(b - a, x.shape[1], -1)
        out.append(c.transpose(0, 2, 1).ravel())
    return np.ascontiguousarray(np.concatenate(out))


def config4_mixed_split(nx: int, ny: int, seed: int = SEED) -> np.ndarray:
    """The seeded third of the cells that configuration 4 cuts into two TRIs (100 k TRI + 100 k QUAD on 500 x 300)."""
    r = SplitMix64(seed + 40).uniform(nx * ny, 0.0, 1.0)
    return r < 1.0 / 3.0


def config4_mixed_mesh(nx: int = 500, ny: int = 300, seed: int = SEED) -> MixedMesh:
    """Configuration 4's mesh: the 500 x 300 periodic base grid of [0,1]^2 with a seeded third of its cells cut into
    two TRIs, orders p ~ U{1..5} per element (seeded)."""
    split = config4_mixed_split(nx, ny, seed)
    n_el = int(nx * ny + split.sum())
    r = SplitMix64(seed + 41).uniform(n_el, 0.0, 1.0)
    p = (1 + np.minimum((r * 5).astype(np.int32), 4)).astype(np.int32)
    return structured_mixed_mesh((nx, ny), split, p)


def config4_mixed_state(mesh: MixedMesh, seed: int = SEED) -> np.ndarray:
    return mixed_state(mesh, lambda x: prim_to_cons(smooth_state_3d(x, seed + 1), 2))
