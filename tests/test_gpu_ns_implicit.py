"""GPU parity of the Navier-Stokes (BR2) linearisation and of its implicit steady solve.

* J v (dgx_jacobian_vector: dual numbers through the gradient-trace and residual kernels, the BR2 gradient
  differentiated with the rest) equals the oracle's complex step of the BR2 residual (or_compute_rhs_ns_c) to
  1e-12 relative: curved periodic meshes p = 1..5 (Lax-Friedrichs and Roe), the Taylor-Couette annulus with its
  rotating, adiabatic and diabatic no-slip walls, and a Riemann box.
* The diagonal blocks dR_e/du_e (dgx_assemble_diag_blocks: one dual-number column per launch, the neighbours'
  states and gradient traces held fixed) equal the oracle's complex-step blocks to 1e-12, with and without the
  -s_e M_e mass term.
* dgx_newton_operator (J v - s_e M_e v in one launch) equals dgx_jacobian_vector followed by dgx_mass_sub.
* The face blocks dR_e/du_n (dgx_assemble_face_blocks: full nb x nb, the neighbour's whole state entering through
  its lifted gradient) equal the oracle's complex-step blocks to 1e-12, and with the diagonal blocks they are the
  whole Jacobian: D_e v_e + sum_faces B_e,n v_n = (J v)_e.
* The multicolour block SGS sweeps over those blocks (dgx_precond_apply, FP64 and FP32-stored) equal the same
  sweeps in numpy, and approximate A^-1 v better than block-Jacobi.
* The pseudo-time step takes the viscous bound dt <= cfl h^2 / mu (solve_dg.c:300-318) where it is the smaller.
* A Newton-GMRES steady solve with block-Jacobi reaches max |R| < 1e-10 on the Taylor-Couette annulus; the
  oracle's residual of the converged state agrees, and the steady velocity is the Couette profile.
* Block SGS reaches the same steady state as block-Jacobi in fewer GMRES iterations.
* The learned preconditioner (trained on Euler blocks) and the FP32 J v are refused for Navier-Stokes with
  DGX_ERR_UNSUPPORTED.
"""
import numpy as np
import pytest

import pyoracle as O
from dpgsolver_amd import mesh as M

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

MU, PR = 0.02, 0.72


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def _ctx(mesh, flux, walls=None, bc_prim=None, mu=MU):
    from dpgsolver_amd import _lib

    ctx = _lib.Context(2, mesh.p, M.PDE_NAVIER_STOKES, flux, mu=mu, Pr=PR, walls=walls, bc_prim=bc_prim)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    ctx.upload_mesh(mesh.xyz, mesh.faces)
    return ctx


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _perturbed(mesh, u, amp=0.02):
    x = M.nodes_xyz(mesh)
    pert = amp * np.sin(3 * x[..., 0]) * np.cos(2 * x[..., 1])
    return u * (1.0 + M.pack_state(np.stack([pert, 2 * pert, -pert, 0.5 * pert], -1)))


def _cases():
    curve = lambda pts: np.stack([pts[..., 0] + 0.03 * np.sin(2 * np.pi * pts[..., 1]),
                                  pts[..., 1] + 0.03 * np.sin(2 * np.pi * pts[..., 0])], -1)
    out = []
    for p, n in ((1, 4), (2, 3), (3, 3), (4, 2), (5, 2)):
        for flux in (M.FLUX_LF, M.FLUX_ROE):
            mesh = M.structured_tp_mesh(2, p, (n, n), (0.0, 0.0), (1.0, 1.0), (True, True), curve=curve,
                                        rotate=np.random.default_rng(p).integers(0, 4, n * n))
            x = M.nodes_xyz(mesh)
            X, Y = x[..., 0], x[..., 1]
            prim = np.stack([1 + 0.1 * np.sin(2 * np.pi * X), 0.3 + 0.2 * np.sin(2 * np.pi * Y),
                             -0.1 + 0.2 * np.cos(2 * np.pi * X), 1 + 0.1 * np.sin(2 * np.pi * (X + Y))], -1)
            out.append((f"periodic_p{p}_{'lf' if flux == M.FLUX_LF else 'roe'}", mesh, flux, {},
                        M.pack_state(M.prim_to_cons(prim, 2))))
    for p, bc_outer, dT in ((2, M.BC_NOSLIP_ADIABATIC, 0.0), (3, M.BC_NOSLIP_DIABATIC, 1.0)):
        mesh = M.annulus_mesh(p, 2, 8, bc_outer=bc_outer)
        u, _ = M.taylor_couette_state(mesh, MU, PR, dTdr_o=dT)
        out.append((f"annulus_p{p}_{'diabatic' if dT else 'adiabatic'}", mesh, M.FLUX_ROE,
                    {"walls": M.taylor_couette_walls(MU, PR, dTdr_o=dT)}, _perturbed(mesh, u)))
    prim = np.array([1.0, 0.3, 0.2, 1.0])
    mesh = M.structured_tp_mesh(2, 2, (3, 3), (0.0, 0.0), (1.0, 0.8), (False, False),
                                bc_low=(M.BC_RIEMANN, M.BC_RIEMANN), bc_high=(M.BC_RIEMANN, M.BC_RIEMANN))
    out.append(("riemann_box_p2", mesh, M.FLUX_LF, {"bc_prim": prim},
                _perturbed(mesh, M.uniform_state(mesh, prim))))
    return out


CASES = _cases()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_ns_jacobian_vector_matches_complex_step(dev, case):
    name, mesh, flux, kw, u = case
    print(name, flush=True)
    ctx = _ctx(mesh, flux, **kw)
    v = np.random.default_rng(9).standard_normal(u.size) * np.abs(u)
    ud, vd = torch.from_numpy(u).to(dev), torch.from_numpy(v).to(dev)
    jv = torch.full_like(ud, float("nan"))
    ctx.jacobian_vector(ud.data_ptr(), vd.data_ptr(), jv.data_ptr())
    torch.cuda.synchronize()
    ref = O.Oracle(mesh, M.PDE_EULER, flux, viscous=(MU, PR), **kw).jvp_ns(u, v)
    err = _rel(jv.cpu().numpy(), ref)
    print(f"  J v rel err {err:.1e}", flush=True)
    assert err < 1e-12
    # J v - s_e M_e v in one launch
    sc = np.random.default_rng(19).uniform(0.5, 30.0, mesh.E)
    sd = torch.from_numpy(sc).to(dev)
    a = torch.full_like(ud, float("nan"))
    ctx.newton_operator(ud.data_ptr(), sd.data_ptr(), vd.data_ptr(), a.data_ptr())
    ctx.mass_sub(vd.data_ptr(), sd.data_ptr(), jv.data_ptr())
    torch.cuda.synchronize()
    assert _rel(a.cpu().numpy(), jv.cpu().numpy()) < 1e-13
    ctx.close()


BLOCK_CASES = [c for c in CASES if c[0] in ("periodic_p2_roe", "periodic_p3_lf", "annulus_p2_adiabatic",
                                             "annulus_p3_diabatic", "riemann_box_p2")]


@pytest.mark.parametrize("case", BLOCK_CASES, ids=[c[0] for c in BLOCK_CASES])
def test_ns_diag_blocks_match_complex_step(dev, case):
    name, mesh, flux, kw, u = case
    print(name, flush=True)
    ctx = _ctx(mesh, flux, **kw)
    nb = ctx.Np * ctx.nv
    ud = torch.from_numpy(u).to(dev)
    blocks = torch.full((mesh.E * nb * nb,), float("nan"), dtype=torch.float64, device=dev)
    ctx.assemble_diag_blocks(ud.data_ptr(), None, blocks.data_ptr())
    torch.cuda.synchronize()
    g = blocks.cpu().numpy().reshape(mesh.E, nb, nb)
    orc = O.Oracle(mesh, M.PDE_EULER, flux, viscous=(MU, PR), **kw)
    s = np.linspace(0.5, 2.0, mesh.E)
    sd = torch.from_numpy(s).to(dev)
    ctx.assemble_diag_blocks(ud.data_ptr(), sd.data_ptr(), blocks.data_ptr())
    torch.cuda.synchronize()
    g2 = blocks.cpu().numpy().reshape(mesh.E, nb, nb)
    for e in sorted({0, mesh.E // 2, mesh.E - 1}):
        ref = orc.ns_diag_block(u, e)
        err = np.abs(g[e] - ref).max() / np.abs(ref).max()
        print(f"  element {e}: block rel err {err:.1e}", flush=True)
        assert err < 1e-12
        Mm = np.kron(np.eye(ctx.nv), orc.mass(e))
        assert np.abs(g2[e] - (ref - s[e] * Mm)).max() <= 1e-12 * np.abs(ref).max()
    ctx.close()


def _neighbours(mesh):
    """(e, lf) -> neighbour element, from the face table rows (eL, fL, eR, fR, ord, bc)."""
    nb = {}
    for eL, fL, eR, fR, _, _ in mesh.faces:
        if eR >= 0:
            nb[(int(eL), int(fL))] = int(eR)
            nb[(int(eR), int(fR))] = int(eL)
    return nb


FACE_CASES = [c for c in CASES if c[0] in ("periodic_p1_roe", "periodic_p2_lf", "annulus_p2_adiabatic",
                                            "riemann_box_p2")]


def _periodic_p4_3x3():
    """p = 4 (nb = 100, more than 64 rows per block in the full-block SGS sweep) on a 3 x 3 periodic mesh, so that
    every face has its own neighbour."""
    curve = lambda pts: np.stack([pts[..., 0] + 0.03 * np.sin(2 * np.pi * pts[..., 1]),
                                  pts[..., 1] + 0.03 * np.sin(2 * np.pi * pts[..., 0])], -1)
    mesh = M.structured_tp_mesh(2, 4, (3, 3), (0.0, 0.0), (1.0, 1.0), (True, True), curve=curve,
                                rotate=np.random.default_rng(44).integers(0, 4, 9))
    x = M.nodes_xyz(mesh)
    X, Y = x[..., 0], x[..., 1]
    prim = np.stack([1 + 0.1 * np.sin(2 * np.pi * X), 0.3 + 0.2 * np.sin(2 * np.pi * Y),
                     -0.1 + 0.2 * np.cos(2 * np.pi * X), 1 + 0.1 * np.sin(2 * np.pi * (X + Y))], -1)
    return ("periodic_3x3_p4_roe", mesh, M.FLUX_ROE, {}, M.pack_state(M.prim_to_cons(prim, 2)))


FACE_CASES = FACE_CASES + [_periodic_p4_3x3()]


@pytest.mark.parametrize("case", FACE_CASES, ids=[c[0] for c in FACE_CASES])
def test_ns_face_blocks_match_complex_step(dev, case):
    name, mesh, flux, kw, u = case
    print(name, flush=True)
    ctx = _ctx(mesh, flux, **kw)
    nb = ctx.Np * ctx.nv
    ud = torch.from_numpy(u).to(dev)
    fb = torch.full((mesh.E * 4 * nb * nb,), float("nan"), dtype=torch.float64, device=dev)
    db = torch.full((mesh.E * nb * nb,), float("nan"), dtype=torch.float64, device=dev)
    ctx.assemble_face_blocks(ud.data_ptr(), fb.data_ptr())
    ctx.assemble_diag_blocks(ud.data_ptr(), None, db.data_ptr())
    torch.cuda.synchronize()
    F = fb.cpu().numpy().reshape(mesh.E, 4, nb, nb)
    D = db.cpu().numpy().reshape(mesh.E, nb, nb)
    orc = O.Oracle(mesh, M.PDE_EULER, flux, viscous=(MU, PR), **kw)
    nbr = _neighbours(mesh)
    for e in sorted({0, mesh.E - 1}):
        for lf in range(4):
            ref = orc.ns_face_block(u, e, lf)
            scale = max(np.abs(ref).max(), np.abs(D[e]).max())
            err = np.abs(F[e, lf] - ref).max() / scale
            print(f"  element {e} face {lf}: block rel err {err:.1e}", flush=True)
            assert err < 1e-12
            if (e, lf) not in nbr:
                assert not F[e, lf].any()
    # the diagonal and face blocks are the whole Jacobian
    v = np.random.default_rng(5).standard_normal(u.size)
    V = v.reshape(mesh.E, nb)
    Jv = V.copy()
    for e in range(mesh.E):
        Jv[e] = D[e] @ V[e] + sum(F[e, lf] @ V[n] for (ee, lf), n in nbr.items() if ee == e)
    ref = orc.jvp_ns(u, v)
    assert _rel(Jv.ravel(), ref) < 1e-12
    ctx.close()


@pytest.mark.parametrize("case", FACE_CASES[2:], ids=[c[0] for c in FACE_CASES[2:]])
def test_ns_jacobian_bsr_is_the_newton_operator(dev, case):
    name, mesh, flux, kw, u = case
    ctx = _ctx(mesh, flux, **kw)
    nb = ctx.Np * ctx.nv
    rp, ci = ctx.jacobian_bsr_pattern()
    ud = torch.from_numpy(u).to(dev)
    sc = np.random.default_rng(3).uniform(0.5, 5.0, mesh.E)
    sd = torch.from_numpy(sc).to(dev)
    vals = torch.full((len(ci) * nb * nb,), float("nan"), dtype=torch.float64, device=dev)
    ctx.assemble_jacobian_bsr(ud.data_ptr(), sd.data_ptr(), vals.data_ptr())
    v = np.random.default_rng(4).standard_normal(u.size)
    vd = torch.from_numpy(v).to(dev)
    ref = torch.empty_like(ud)
    ctx.newton_operator(ud.data_ptr(), sd.data_ptr(), vd.data_ptr(), ref.data_ptr())
    torch.cuda.synchronize()
    W = vals.cpu().numpy().reshape(-1, nb, nb)
    V = v.reshape(mesh.E, nb)
    got = np.zeros_like(V)
    for e in range(mesh.E):
        for k in range(rp[e], rp[e + 1]):
            got[e] += W[k] @ V[ci[k]]
    assert _rel(got.ravel(), ref.cpu().numpy()) < 1e-12
    ctx.close()


def _greedy_colours(mesh, nbr):
    """Smallest colour unused by the already coloured face neighbours, elements in index order."""
    col = [-1] * mesh.E
    for e in range(mesh.E):
        used = {col[n] for (ee, lf), n in nbr.items() if ee == e and n != e and col[n] >= 0}
        k = 0
        while k in used:
            k += 1
        col[e] = k
    return col


@pytest.mark.parametrize("case", FACE_CASES, ids=[c[0] for c in FACE_CASES])
def test_ns_block_sgs_apply_matches_numpy(dev, case):
    from dpgsolver_amd import _lib

    name, mesh, flux, kw, u = case
    ctx = _ctx(mesh, flux, **kw)
    nb = ctx.Np * ctx.nv
    ud = torch.from_numpy(u).to(dev)
    inv_dt = torch.empty(mesh.E, dtype=torch.float64, device=dev)
    ctx.cfl_dt(ud.data_ptr(), 5.0, inv_dt.data_ptr())
    fb = torch.zeros(mesh.E * 4 * nb * nb, dtype=torch.float64, device=dev)
    db = torch.zeros(mesh.E * nb * nb, dtype=torch.float64, device=dev)
    ctx.assemble_face_blocks(ud.data_ptr(), fb.data_ptr())
    ctx.assemble_diag_blocks(ud.data_ptr(), inv_dt.data_ptr(), db.data_ptr())
    v = np.random.default_rng(23).standard_normal(u.size)
    vd = torch.from_numpy(v).to(dev)
    zb, zs, z32 = torch.empty_like(ud), torch.empty_like(ud), torch.empty_like(ud)
    ctx.precond_apply(_lib.PC_BLOCK_JACOBI, ud.data_ptr(), inv_dt.data_ptr(), vd.data_ptr(), zb.data_ptr())
    ctx.precond_apply(_lib.PC_BLOCK_SGS, ud.data_ptr(), inv_dt.data_ptr(), vd.data_ptr(), zs.data_ptr())
    ctx.precond_apply(_lib.PC_BLOCK_SGS32, ud.data_ptr(), inv_dt.data_ptr(), vd.data_ptr(), z32.data_ptr())
    torch.cuda.synchronize()
    F = fb.cpu().numpy().reshape(mesh.E, 4, nb, nb)
    D = db.cpu().numpy().reshape(mesh.E, nb, nb)
    V = v.reshape(mesh.E, nb)
    nbr = _neighbours(mesh)
    assert _rel(zb.cpu().numpy(), np.linalg.solve(D, V[..., None])[..., 0].ravel()) < 1e-12
    col = _greedy_colours(mesh, nbr)
    colours = sorted(set(col))
    z = np.zeros_like(V)
    for sweep in (colours, colours[::-1]):       # forward, then backward
        for k in sweep:
            zold = z.copy()
            for e in (e for e in range(mesh.E) if col[e] == k):
                r = V[e].copy()
                for lf in range(4):
                    if (e, lf) in nbr:
                        r -= F[e, lf] @ zold[nbr[(e, lf)]]
                z[e] = np.linalg.solve(D[e], r)
    got = zs.cpu().numpy()
    err = _rel(got, z.ravel())
    e32 = _rel(z32.cpu().numpy(), z.ravel())
    print(f"{name}: SGS vs numpy {err:.1e}, FP32-stored {e32:.1e}", flush=True)
    assert err < 1e-11 and 0 < e32 < 1e-5
    A = np.zeros((mesh.E * nb, mesh.E * nb))
    for e in range(mesh.E):
        A[e * nb:(e + 1) * nb, e * nb:(e + 1) * nb] += D[e]
    for (e, lf), n in nbr.items():
        A[e * nb:(e + 1) * nb, n * nb:(n + 1) * nb] += F[e, lf]
    x = np.linalg.solve(A, v)
    assert np.linalg.norm(got - x) < np.linalg.norm(zb.cpu().numpy() - x)
    ctx.close()


@pytest.mark.parametrize("flux", [M.FLUX_ROE, M.FLUX_LF])
def test_ns_sutherland_residual_and_linearisation_match_oracle(dev, flux):
    """Sutherland's viscosity (DGX_VISC_SUTHERLAND, r_s 0.4): the FP64 residual, J v and the diagonal blocks equal
    the oracle to 1e-12, on a curved periodic mesh and on the no-slip annulus; the mixed (perturbation) residual
    stays within 1e-5 of FP64; variable-order and mixed TRI/QUAD meshes refuse it."""
    from dpgsolver_amd import _lib

    for name, mesh, fl, kw, u in (CASES[2], CASES[10]):
        # the annulus keeps its Roe flux: the outer wall's nodes are at rest, where the Lax-Friedrichs max(|V| + c)
        # has no derivative (|V| at V = 0; the reference's dmaxV_ds divides by V there too,
        # numerical_flux_euler_T.c:1653-1676)
        flux = flux if name.startswith("periodic") else fl
        ctx = _lib.Context(2, mesh.p, M.PDE_NAVIER_STOKES, flux, mu=MU, Pr=PR, sutherland=True, r_s=0.4, **kw)
        ctx.set_stream(torch.cuda.current_stream().cuda_stream)
        ctx.upload_mesh(mesh.xyz, mesh.faces)
        orc = O.Oracle(mesh, M.PDE_EULER, flux, viscous=(MU, PR), sutherland_r_s=0.4, **kw)
        ud = torch.from_numpy(u).to(dev)
        r = torch.full_like(ud, float("nan"))
        ctx.compute_rhs(ud.data_ptr(), r.data_ptr())
        v = np.random.default_rng(2).standard_normal(u.size) * np.abs(u)
        vd = torch.from_numpy(v).to(dev)
        jv = torch.full_like(ud, float("nan"))
        ctx.jacobian_vector(ud.data_ptr(), vd.data_ptr(), jv.data_ptr())
        rmx = torch.full_like(ud, float("nan"), dtype=torch.float32)
        ctx.compute_rhs_mixed(ud.data_ptr(), rmx.data_ptr())
        nb = ctx.Np * ctx.nv
        blocks = torch.full((mesh.E * nb * nb,), float("nan"), dtype=torch.float64, device=dev)
        ctx.assemble_diag_blocks(ud.data_ptr(), None, blocks.data_ptr())
        torch.cuda.synchronize()
        ref = orc.rhs_ns(u)
        e_r = _rel(r.cpu().numpy(), ref)
        e_jv = _rel(jv.cpu().numpy(), orc.jvp_ns(u, v))
        e_mx = _rel(rmx.double().cpu().numpy(), ref)
        blk = orc.ns_diag_block(u, 0)
        e_b = np.abs(blocks.cpu().numpy().reshape(mesh.E, nb, nb)[0] - blk).max() / np.abs(blk).max()
        print(f"{name} flux {flux}: residual {e_r:.1e}, J v {e_jv:.1e}, block {e_b:.1e}, mixed {e_mx:.1e}", flush=True)
        assert e_r < 1e-12 and e_jv < 1e-12 and e_b < 1e-12 and e_mx < 1e-5
        ctx.close()
    ctx = _lib.Context(2, 2, M.PDE_NAVIER_STOKES, flux, mu=MU, Pr=PR, sutherland=True)
    hm = M.structured_hp_mesh(np.array([1, 2, 2, 1]), (2, 2), (0.0, 0.0), (1.0, 1.0), (True, True))
    with pytest.raises(_lib.DgxError, match="UNSUPPORTED"):
        ctx.upload_mesh_hp(hm)
    ctx.close()


def test_ns_cfl_dt_takes_the_viscous_bound(dev):
    mesh = M.structured_tp_mesh(2, 2, (4, 4), (0.0, 0.0), (1.0, 1.0), (True, True))
    prim = np.array([1.0, 0.0, 0.0, 1.0])
    u = M.uniform_state(mesh, prim)
    h, cfl, c = 0.25, 3.0, np.sqrt(1.4)
    for mu in (1e-3, 5.0):   # the acoustic bound, then the viscous one
        ctx = _ctx(mesh, M.FLUX_ROE, mu=mu)
        ud = torch.from_numpy(u).to(dev)
        inv = torch.empty(mesh.E, dtype=torch.float64, device=dev)
        ctx.cfl_dt(ud.data_ptr(), cfl, inv.data_ptr())
        torch.cuda.synchronize()
        want = max(c / (cfl * h), mu / (cfl * h * h))
        np.testing.assert_allclose(inv.cpu().numpy(), want, rtol=1e-13)
        ctx.close()


def test_ns_steady_taylor_couette_newton_gmres(dev):
    from dpgsolver_amd import _lib

    mu = 0.1
    walls = M.taylor_couette_walls(mu, PR)
    mesh = M.annulus_mesh(2, 3, 12)
    u0, _ = M.taylor_couette_state(mesh, mu, PR)
    ctx = _ctx(mesh, M.FLUX_ROE, walls=walls, mu=mu)
    ud = torch.from_numpy(_perturbed(mesh, u0, 0.05)).to(dev)
    st = ctx.solve_implicit(ud.data_ptr(), cfl_initial=5.0, exit_tol=1e-10, max_newton=100,
                            precond=_lib.PC_BLOCK_JACOBI)
    print("newton", st["newton_steps"], "gmres", st["gmres_iters"], "max_rhs", st["max_rhs"], st["seconds"],
          flush=True)
    assert st["converged"] == 1 and st["max_rhs"] < 1e-10
    u = ud.cpu().numpy()
    r = O.Oracle(mesh, M.PDE_EULER, M.FLUX_ROE, viscous=(mu, PR), walls=walls).rhs_ns(u)
    assert np.abs(r).max() < 1e-9          # the oracle agrees that this is the steady state
    # the steady velocity is the Couette profile V_theta = c (1/r - r/r_o^2); the pressure takes up rho V^2 / r
    x = M.nodes_xyz(mesh)
    cons = u.reshape(mesh.E, 4, -1).transpose(0, 2, 1)
    X, Y = x[..., 0], x[..., 1]
    R = np.hypot(X, Y)
    vt = (-Y * cons[..., 1] + X * cons[..., 2]) / (R * cons[..., 0])
    vr = (X * cons[..., 1] + Y * cons[..., 2]) / (R * cons[..., 0])
    c = 0.5 / (1.0 / 0.5 - 0.5)
    exact = c * (1.0 / R - R)
    print(f"V_theta error {np.abs(vt - exact).max():.2e}, max |V_r| {np.abs(vr).max():.2e}", flush=True)
    assert np.abs(vt - exact).max() < 0.02 * exact.max() and np.abs(vr).max() < 0.02 * exact.max()
    # block SGS over the full BR2 face blocks: the same steady state in fewer GMRES iterations
    us = torch.from_numpy(_perturbed(mesh, u0, 0.05)).to(dev)
    ss = ctx.solve_implicit(us.data_ptr(), cfl_initial=5.0, exit_tol=1e-10, max_newton=100,
                            precond=_lib.PC_BLOCK_SGS)
    print("block SGS: newton", ss["newton_steps"], "gmres", ss["gmres_iters"], ss["seconds"], flush=True)
    assert ss["converged"] == 1 and ss["gmres_iters"] < st["gmres_iters"]
    assert _rel(us.cpu().numpy(), u) < 1e-8
    # the GMRES operator in FP32 dual numbers (the BR2 residual's J v, gradients included), the FP64 Newton
    # residual deciding convergence: the same steady state, the same Newton steps, GMRES iterations within 5%
    u32 = torch.from_numpy(_perturbed(mesh, u0, 0.05)).to(dev)
    s32 = ctx.solve_implicit(u32.data_ptr(), cfl_initial=5.0, exit_tol=1e-10, max_newton=100,
                             precond=_lib.PC_BLOCK_SGS, jvp_fp32=True)
    d = _rel(u32.cpu().numpy(), us.cpu().numpy())
    print("block SGS, FP32 operator: newton", s32["newton_steps"], "gmres", s32["gmres_iters"], s32["seconds"],
          f"state rel diff {d:.2e}", flush=True)
    assert s32["converged"] == 1 and s32["max_rhs"] < 1e-10 and d < 1e-9
    assert s32["newton_steps"] == ss["newton_steps"]
    assert abs(s32["gmres_iters"] - ss["gmres_iters"]) <= 0.05 * ss["gmres_iters"] + 4
    ctx.close()


def test_ns_learned_preconditioner_is_refused(dev):
    from dpgsolver_amd import _lib

    mesh = M.structured_tp_mesh(2, 2, (3, 3), (0.0, 0.0), (1.0, 1.0), (True, True))
    ctx = _ctx(mesh, M.FLUX_ROE)
    u = torch.from_numpy(M.uniform_state(mesh, np.array([1.0, 0.3, 0.1, 1.0]))).to(dev)
    with pytest.raises(_lib.DgxError, match="UNSUPPORTED"):
        ctx.solve_implicit(u.data_ptr(), precond=_lib.PC_LEARNED)
    ctx.close()
