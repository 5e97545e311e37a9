"""Pins the TRI / QUAD mixed-mesh oracle (oracle/mixed_oracle.py) before it checks the GPU path.

* TRI node sets: the oracle's and the library's expansion of the reference's orbit tables equal the golden
  Nodes_d2_p3_WSH and Nodes_d2_p6_WV of input/testing/unit/nodes/nodes_si.data; the library's reference nodes are
  the oracle's.
* WV cubature of degree q integrates every monomial of degree <= q over the reference triangle exactly (against
  a collapsed Gauss-Legendre product rule), and fails one degree higher.
* On QUAD-only meshes the mixed oracle equals the variable-order oracle (hp_oracle, itself pinned against the C
  oracle), curved and with random orders.
* On mixed meshes: free-stream preservation (straight elements), discrete conservation, and M^-1 R converging to
  the exact -div F of a smooth state at order >= p - 0.5.
"""
import json
import math
import os

import numpy as np
import pytest

import hp_oracle as HO
import mixed_oracle as MO
from dpgsolver_amd import mesh as M

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "tri_nodes.json")))


@pytest.mark.parametrize("name,kind,q", [("Nodes_d2_p3_WSH", "WSH", 3), ("Nodes_d2_p6_WV", "WV", 6)])
def test_tri_nodes_match_reference_golden(name, kind, q):
    g = GOLD["golden"][name]
    _, rst, w = MO.tri_nodes(kind, q)
    np.testing.assert_allclose(rst, np.array(g["rst"]), atol=2e-15, rtol=0)
    np.testing.assert_allclose(w, np.array(g["w"]), atol=2e-15, rtol=0)
    from dpgsolver_amd import _lib

    lrst, lw = _lib.tri_rule(0 if kind == "WSH" else 1, q)
    np.testing.assert_allclose(lrst, np.array(g["rst"]), atol=2e-15, rtol=0)
    np.testing.assert_allclose(lw, np.array(g["w"]), atol=2e-15, rtol=0)


@pytest.mark.parametrize("p", range(1, 6))
def test_library_reference_nodes_are_the_oracles(p):
    from dpgsolver_amd import _lib

    rst, w = _lib.reference_nodes(1, p)
    _, ors, ow = MO.tri_nodes("WSH", p)
    np.testing.assert_allclose(rst, ors, atol=5e-16, rtol=0)       # the same expansion, summed in another order
    np.testing.assert_array_equal(w, ow)
    assert len(w) == (p + 1) * (p + 2) // 2
    assert abs(w.sum() - math.sqrt(3.0)) < 1e-14          # area of the reference triangle


def _exact_monomial(a, b):
    """Integral of r^a s^b over the reference triangle by a collapsed 20 x 20 Gauss-Legendre rule (exact here)."""
    x, w = np.polynomial.legendre.leggauss(20)
    u, wu = 0.5 * (x + 1), 0.5 * w
    V0, V1, V2 = MO.TRI_V
    e1, e2 = V1 - V0, V2 - V0
    det = abs(e1[0] * e2[1] - e1[1] * e2[0])
    tot = 0.0
    for i in range(20):
        for j in range(20):
            xi, eta = u[i], u[j] * (1 - u[i])
            pt = V0 + xi * (V1 - V0) + eta * (V2 - V0)
            tot += wu[i] * wu[j] * (1 - u[i]) * det * pt[0] ** a * pt[1] ** b
    return tot


@pytest.mark.parametrize("q", [2, 4, 6, 8, 10])
def test_wv_cubature_is_exact_to_its_degree(q):
    _, rst, w = MO.tri_nodes("WV", q)
    worst = 0.0
    for tot in range(q + 1):
        for b in range(tot + 1):
            a = tot - b
            worst = max(worst, abs(w @ (rst[:, 0] ** a * rst[:, 1] ** b) - _exact_monomial(a, b)))
    assert worst < 1e-13, worst
    # and not beyond it (pins the degree, i.e. the WV<2p> choice)
    miss = max(abs(w @ (rst[:, 0] ** a * rst[:, 1] ** (q + 1 + k - a)) - _exact_monomial(a, q + 1 + k - a))
               for k in (0, 1) for a in range(q + 2 + k))
    assert miss > 1e-10


def _curve(amp=0.03):
    def f(pts):
        out = pts.copy()
        out[..., 0] = pts[..., 0] + amp * np.sin(2 * np.pi * pts[..., 1])
        out[..., 1] = pts[..., 1] + amp * np.sin(2 * np.pi * pts[..., 0])
        return out
    return f


def test_mixed_oracle_equals_hp_oracle_on_quad_meshes():
    rng = np.random.default_rng(31)
    n = (4, 3)
    p = rng.integers(1, 6, 12).astype(np.int32)
    mm = M.structured_mixed_mesh(n, np.zeros(12, bool), p, curve=_curve())
    hm = M.structured_hp_mesh(p, n, (0.0, 0.0), (1.0, 1.0), (True, True), curve=_curve())
    np.testing.assert_array_equal(mm.p, hm.p)
    for e in range(mm.E):
        np.testing.assert_allclose(mm.xyz[e], hm.xyz[e], atol=1e-15)
    state = lambda x: M.prim_to_cons(M.smooth_state_3d(x), 2)
    um = M.mixed_state(mm, state)
    uh = M.hp_state(hm, state)
    np.testing.assert_allclose(um, uh, atol=1e-15)
    rm = MO.MixedOracle(mm, M.FLUX_LF).rhs(um)
    rh = HO.HpOracle(hm, 1, M.FLUX_LF).rhs(uh)
    assert np.abs(rm - rh).max() < 1e-13 * np.abs(rh).max()


def _mixed(n=(4, 4), p=None, seed=5, curve=None, split=None):
    rng = np.random.default_rng(seed)
    if split is None:
        split = rng.random(n[0] * n[1]) < 0.5
    n_el = int(n[0] * n[1] + np.sum(split))
    pe = rng.integers(1, 6, n_el).astype(np.int32) if p is None else np.full(n_el, p, np.int32)
    return M.structured_mixed_mesh(n, split, pe, curve=curve)


@pytest.mark.parametrize("flux", [M.FLUX_LF, M.FLUX_ROE])
@pytest.mark.parametrize("orders", ["uniform3", "random"])
def test_free_stream_and_conservation_on_mixed_meshes(flux, orders):
    mesh = _mixed(p=3 if orders == "uniform3" else None)
    assert set(mesh.kind) == {0, 1}
    orc = MO.MixedOracle(mesh, flux)
    prim = np.array([1.0, 0.3, -0.2, 1.0])
    u = M.mixed_state(mesh, lambda x: np.broadcast_to(M.prim_to_cons(prim, 2), (len(x), 4)))
    r = orc.rhs(u)
    assert np.abs(r).max() < 1e-13, np.abs(r).max()
    # conservation: the Lagrange basis sums to one, so the nodal residual sums to the boundary flux (zero here)
    u2 = M.mixed_state(mesh, lambda x: M.prim_to_cons(M.smooth_state_3d(x), 2))
    r2 = orc.rhs(u2)
    off = orc.off
    tot = sum(r2[off[e]:off[e + 1]].reshape(4, -1).sum(1) for e in range(mesh.E))
    assert np.abs(tot).max() < 5e1 * 1e-15 * mesh.E, tot
    assert np.abs(r2).max() > 1e-3


def _smooth_prim(x):
    X, Y = x[:, 0], x[:, 1]
    rho = 1.0 + 0.1 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y)
    u = 0.3 + 0.05 * np.cos(2 * np.pi * X)
    v = -0.2 + 0.05 * np.sin(2 * np.pi * Y)
    p = 1.0 + 0.1 * np.cos(2 * np.pi * (X + Y))
    return np.stack([rho, u, v, p], 1)


def _exact_div(x, h=1e-5):
    """-div F of the smooth state at points x (central differences of the exact flux)."""
    def F(pts, k):
        c = M.prim_to_cons(_smooth_prim(pts), 2)
        out = np.empty((len(pts), 4))
        for i, s in enumerate(c):
            rho, m1, m2, E = s
            u, v = m1 / rho, m2 / rho
            p = 0.4 * (E - 0.5 * rho * (u * u + v * v))
            f = np.array([[m1, m2], [m1 * u + p, m1 * v], [m2 * u, m2 * v + p], [(E + p) * u, (E + p) * v]])
            out[i] = f[:, k]
        return out
    dx = (F(x + [h, 0], 0) - F(x - [h, 0], 0)) / (2 * h)
    dy = (F(x + [0, h], 1) - F(x - [0, h], 1)) / (2 * h)
    return -(dx + dy)


@pytest.mark.parametrize("p", [2, 3])
@pytest.mark.parametrize("curved", [False, True])
def test_mixed_residual_converges_to_the_divergence(p, curved):
    """M^-1 R - (-div F) at the nodes is the lift of face jumps of size h^(p+1): order p (measured 2.1-2.8 from
    8^2 to 16^2 cells; the bound leaves 0.5 for the pre-asymptotic range)."""
    errs = []
    for n in (8, 16):
        split = (np.add.outer(np.arange(n), np.arange(n)) % 2 == 0).ravel()     # checkerboard of TRI pairs
        mesh = M.structured_mixed_mesh((n, n), split, np.full(n * n + split.sum(), p, np.int32),
                                       curve=_curve(0.02) if curved else None)
        orc = MO.MixedOracle(mesh, M.FLUX_ROE)
        u = M.mixed_state(mesh, lambda x: M.prim_to_cons(_smooth_prim(x), 2))
        du = orc.apply_minv(orc.rhs(u))
        worst = 0.0
        for e in range(mesh.E):
            ex = _exact_div(mesh.xyz[e].T)
            worst = max(worst, np.abs(du[orc.off[e]:orc.off[e + 1]].reshape(4, -1).T - ex).max())
        errs.append(worst)
    rate = math.log2(errs[0] / errs[1])
    assert rate >= p - 0.5, (errs, rate)


# ---------------------------------------------------------------------------------------------- Navier-Stokes
MU, PR = 0.05, 0.72


def test_mixed_ns_oracle_equals_hp_oracle_on_quad_meshes():
    rng = np.random.default_rng(32)
    n = (4, 3)
    p = rng.integers(1, 6, 12).astype(np.int32)
    mm = M.structured_mixed_mesh(n, np.zeros(12, bool), p, curve=_curve())
    hm = M.structured_hp_mesh(p, n, (0.0, 0.0), (1.0, 1.0), (True, True), curve=_curve())
    state = lambda x: M.prim_to_cons(M.smooth_state_3d(x), 2)
    um, uh = M.mixed_state(mm, state), M.hp_state(hm, state)
    # Lax-Friedrichs: the two mesh builders list some faces with left and right swapped, and the reference's
    # Roe-Pike flux with its entropy fix is not invariant under that swap (F(a, b, n) != -F(b, a, -n))
    rm = MO.MixedOracle(mm, M.FLUX_LF, viscous=(MU, PR)).rhs_ns(um)
    rh = HO.HpOracle(hm, 1, M.FLUX_LF, viscous=(MU, PR)).rhs_ns(uh)
    assert np.abs(rm - rh).max() < 1e-13 * np.abs(rh).max()
    assert np.abs(rh - HO.HpOracle(hm, 1, M.FLUX_LF).rhs(uh)).max() > 1e-3 * np.abs(rh).max()


def test_mixed_ns_constants_and_conservation():
    prim = np.array([1.1, 0.3, -0.2, 0.9])
    const = lambda x: np.broadcast_to(M.prim_to_cons(prim, 2), (len(x), 4))
    straight = _mixed(seed=8)
    orc = MO.MixedOracle(straight, M.FLUX_ROE, viscous=(MU, PR))
    assert np.abs(orc.rhs_ns(M.mixed_state(straight, const))).max() < 1e-13
    curved = _mixed(seed=9, curve=_curve())
    orc = MO.MixedOracle(curved, M.FLUX_LF, viscous=(MU, PR))
    r = orc.rhs_ns(M.mixed_state(curved, lambda x: M.prim_to_cons(_smooth_prim(x), 2)))
    tot = sum(r[orc.off[e]:orc.off[e + 1]].reshape(4, -1).sum(1) for e in range(curved.E))
    assert np.abs(tot).max() < 1e-14 * np.abs(r).sum()


def test_mixed_ns_gradient_of_a_linear_state_and_its_integral():
    """The BR2 gradient G of u = a + b . x is b exactly in every element with no face on the periodic wrap (the
    state is continuous elsewhere, so every lift there vanishes); and on any periodic mesh the integrals of G
    sum to zero (1^T M_e G_e = int grad u + int_boundary n (u* - u), and the averaged traces u* cancel between
    the two sides), which fails for a lift of the wrong sign or a side's own trace in place of u*."""
    mesh = _mixed(n=(5, 4), seed=10)
    orc = MO.MixedOracle(mesh, M.FLUX_ROE, viscous=(MU, PR))
    b = np.array([[0.3, -0.7], [1.1, 0.2], [-0.4, 0.5], [0.9, 1.3]])          # (nv, d)
    lin = lambda x: np.array([1.0, 0.5, -0.2, 2.0]) + x @ b.T
    G, _, _ = orc.ns_gradient(M.mixed_state(mesh, lin))
    wrap = set()
    for eL, fL, eR, fR, ord_, bc in mesh.faces:
        cL, cR = mesh.xyz[eL].mean(1), mesh.xyz[eR].mean(1)
        if np.abs(cL - cR).max() > 0.5:
            wrap.update((int(eL), int(eR)))
    inner = [e for e in range(mesh.E) if e not in wrap]
    assert len(inner) >= 6 and len(wrap) > 0
    for e in inner:
        np.testing.assert_allclose(G[e], np.broadcast_to(b.T[:, :, None], G[e].shape), atol=1e-12, rtol=0)
    u = M.mixed_state(mesh, lambda x: M.prim_to_cons(_smooth_prim(x), 2))
    G, gv, _ = orc.ns_gradient(u)
    tot = sum(np.einsum("ij,dvj->dv", orc.mass(e), G[e]) for e in range(mesh.E))
    vol = sum(np.abs(np.einsum("ij,dvj->dv", orc.mass(e), gv[e])).sum() for e in range(mesh.E))
    assert np.abs(tot).max() < 1e-14 * vol, (tot, vol)
    assert vol > 1e-2


# ---------------------------------------------------------------------------------------------- no-slip walls
def test_mixed_ns_walls_equal_the_c_oracle_on_quad_annuli():
    """The mixed oracle's wall faces (its own restatement of the boundary states, gradients and energy flux)
    equal the C oracle's, pinned on Taylor-Couette (tests/test_ns_oracle.py), on all-QUAD annuli."""
    import pyoracle as O

    mu, Pr = 0.1, 0.72
    for p in (2, 3):
        for bco, dT in ((M.BC_NOSLIP_ADIABATIC, 0.0), (M.BC_NOSLIP_DIABATIC, 1.0)):
            qm = M.annulus_mesh(p, 2, 8, bc_outer=bco)
            mm = M.annulus_mixed_mesh((2, 8), np.zeros(16, bool), np.full(16, p, np.int32), bc_outer=bco)
            walls = M.taylor_couette_walls(mu, Pr, dTdr_o=dT)
            u, _ = M.taylor_couette_state(qm, mu, Pr, dTdr_o=dT)
            um, _ = M.taylor_couette_mixed_state(mm, mu, Pr, dTdr_o=dT)
            assert np.abs(u - um).max() < 1e-14
            rq = O.Oracle(qm, M.PDE_EULER, M.FLUX_ROE, viscous=(mu, Pr), walls=walls).rhs_ns(u)
            rm = MO.MixedOracle(mm, M.FLUX_ROE, viscous=(mu, Pr), walls=walls).rhs_ns(um)
            # measured 3e-14 to 7e-13 (two constructions of the same operators)
            assert np.abs(rq - rm).max() < 2e-12 * np.abs(rq).max(), (p, bco)


def test_mixed_ns_walls_converge_on_a_mixed_annulus():
    """Taylor-Couette on annuli with half the cells cut into TRIs: M^-1 R converges to the state's -div F
    (measured p = 3: 0.354, then 0.149)."""
    mu, Pr = 0.1, 0.72
    errs = []
    for n in ((2, 8), (4, 16)):
        split = np.random.default_rng(1).random(n[0] * n[1]) < 0.5
        mm = M.annulus_mixed_mesh(n, split, np.full(n[0] * n[1] + split.sum(), 3, np.int32))
        u, rate = M.taylor_couette_mixed_state(mm, mu, Pr)
        orc = MO.MixedOracle(mm, M.FLUX_ROE, viscous=(mu, Pr), walls=M.taylor_couette_walls(mu, Pr))
        errs.append(np.abs(orc.apply_minv(orc.rhs_ns(u)) - rate).max())
    assert errs[0] / errs[1] >= 2.0, errs
