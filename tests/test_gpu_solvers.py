"""GPU parity of the explicit stepper and the implicit path against the CPU oracle.

* SSP-RK3 steps (with the Bezier positivity limiter) against or_ssp_rk3_step.
* Element diagonal Jacobian blocks (forward-mode dual numbers on the GPU) against the oracle's
  complex-step blocks: the reference's own linearization check (analytic vs complex step,
  test_integration_linearization.c:377, 4e-13).
* Exact Jacobian-vector products against complex step of the full oracle residual.
* Batched block LU against numpy; CFL dt against its closed form on a uniform state.
* The block-sparse Newton matrix (the reference's AIJ pattern) applied to v equals the matrix-free operator.
* A steady channel solve to 1e-10, whose converged state is re-checked by the oracle residual.
"""
import numpy as np
import pytest

import pyoracle as O
from dpgsolver_amd import mesh as M

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def _ctx(mesh, pde, flux, b_adv=(0.0, 0.0, 0.0), bc_prim=None):
    from dpgsolver_amd import _lib

    ctx = _lib.Context(mesh.d, mesh.p, pde, flux, b_adv=b_adv, bc_prim=bc_prim)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    ctx.upload_mesh(mesh.xyz, mesh.faces)
    return ctx


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _channel_state(mesh, amp=0.02):
    prim = M.freestream_prim(2)
    x = M.nodes_xyz(mesh)
    pert = amp * np.sin(3 * x[..., :1]) * np.cos(2 * x[..., 1:2]) * np.array([1, 1, 0.5, 1])
    return M.pack_state(M.prim_to_cons(prim[None, None, :] + pert, 2)), prim


# ----------------------------------------------------------------------------- explicit
@pytest.mark.parametrize("case", ["adv", "euler_vortex"])
def test_ssp_rk3_steps_match_oracle(dev, case):
    if case == "adv":
        mesh = M.config1_mesh(8, 2)
        u = M.config1_state(mesh)
        pde, flux, b, dt, pos = M.PDE_ADVECTION, M.FLUX_UPWIND, (1.0, 0.5, 0.0), 0.2 * 0.25 / (5 * 1.0), False
    else:
        mesh = M.config2_mesh(8, 3)
        u = M.config2_state(mesh)
        pde, flux, b, dt, pos = M.PDE_EULER, M.FLUX_ROE, (0.0, 0.0, 0.0), 0.01, True
    ctx = _ctx(mesh, pde, flux, b)
    orc = O.Oracle(mesh, pde, flux, b)
    ud = torch.from_numpy(u).to(dev)
    uo = u.copy()
    for _ in range(5):
        mx = ctx.ssp_rk3_step(ud.data_ptr(), dt, positivity=pos)
        uo, mxo, err = orc.ssp_rk3_step(uo, dt, positivity=pos)
        assert err == 0
        assert abs(mx - mxo) <= 1e-10 * mxo
    torch.cuda.synchronize()
    assert _rel(ud.cpu().numpy(), uo) < 1e-12
    ctx.close()


def test_positivity_limiter_matches_oracle(dev):
    mesh = M.config2_mesh(3, 3)
    u = M.uniform_state(mesh, np.array([1.0, 0.2, 0.1, 1.0])).reshape(mesh.E, 4, 16)
    u[0, 0, 5] = -0.3
    u[4, 3, 7] = -0.5        # negative pressure contribution in another element
    u = u.ravel()
    ctx = _ctx(mesh, M.PDE_EULER, M.FLUX_ROE)
    ud = torch.from_numpy(u).to(dev)
    assert ctx.enforce_positivity(ud.data_ptr()) == 0
    ref, st = O.Oracle(mesh, M.PDE_EULER, M.FLUX_ROE).positivity(u)
    assert st == 0
    assert _rel(ud.cpu().numpy(), ref) < 1e-12
    bad = torch.full_like(ud, -1.0)
    assert ctx.enforce_positivity(bad.data_ptr()) == -4   # DGX_ERR_POSITIVITY, not an abort
    ctx.close()


# ----------------------------------------------------------------------------- implicit building blocks
def _jac_cases():
    out = []
    m = M.structured_tp_mesh(2, 2, (3, 3), (0, 0), (1, 1), (True, True))
    u = M.pack_state(M.prim_to_cons(M.smooth_state_3d(M.nodes_xyz(m)), 2))
    out.append(("periodic_roe_p2", m, M.FLUX_ROE, None, u))
    m = M.structured_tp_mesh(2, 3, (3, 3), (0, 0), (1, 1), (True, True),
                             rotate=np.random.default_rng(2).integers(0, 4, 9))
    u = M.pack_state(M.prim_to_cons(M.smooth_state_3d(M.nodes_xyz(m)), 2))
    out.append(("rotated_lf_p3", m, M.FLUX_LF, None, u))
    m = M.channel_mesh(4, 3, 2)
    u, prim = _channel_state(m)
    out.append(("channel_bc_roe_p2", m, M.FLUX_ROE, prim, u))
    # HEX p = 1 (nb = 40): a curved periodic mesh with rotated element numberings (every face orientation), and a
    # box with Riemann faces on two sides
    curve = lambda pts: pts + 0.03 * np.sin(2 * np.pi * pts[..., [1, 2, 0]])
    m = M.structured_tp_mesh(3, 1, (3, 3, 3), (0, 0, 0), (1, 1, 1), (True, True, True), curve=curve,
                             rotate=np.random.default_rng(4).integers(0, 8, 27))
    u = M.pack_state(M.prim_to_cons(M.smooth_state_3d(M.nodes_xyz(m)), 3))
    out.append(("hex_rotated_roe_p1", m, M.FLUX_ROE, None, u))
    prim = M.freestream_prim(3)
    R = M.BC_RIEMANN
    m = M.structured_tp_mesh(3, 1, (3, 2, 2), (0, 0, 0), (1, 0.75, 0.75), (False, True, True), bc_low=(R, R, R),
                             bc_high=(R, R, R))
    x = M.nodes_xyz(m)
    pert = 0.02 * np.sin(3 * x[..., 0]) * np.cos(2 * x[..., 1]) * np.cos(x[..., 2])
    u = M.pack_state(M.prim_to_cons(prim[None, None, :] * (1.0 + np.stack([pert, pert, -pert, pert, 0.5 * pert], -1)), 3))
    out.append(("hex_box_lf_p1", m, M.FLUX_LF, prim, u))
    # the largest blocks of the in-LDS inverse: QUAD Euler p = 4 (nb = 100) and HEX p = 2 (nb = 135, 270 face dofs)
    m = M.structured_tp_mesh(2, 4, (3, 3), (0, 0), (1, 1), (True, True),
                             rotate=np.random.default_rng(6).integers(0, 4, 9))
    u = M.pack_state(M.prim_to_cons(M.smooth_state_3d(M.nodes_xyz(m)), 2))
    out.append(("rotated_roe_p4", m, M.FLUX_ROE, None, u))
    m = M.structured_tp_mesh(3, 2, (2, 2, 2), (0, 0, 0), (1, 1, 1), (True, True, True), curve=curve,
                             rotate=np.random.default_rng(7).integers(0, 8, 8))
    u = M.pack_state(M.prim_to_cons(M.smooth_state_3d(M.nodes_xyz(m)), 3))
    out.append(("hex_rotated_lf_p2", m, M.FLUX_LF, None, u))
    return out


JAC = _jac_cases()


@pytest.mark.parametrize("case", JAC, ids=[c[0] for c in JAC])
def test_diag_blocks_match_complex_step(dev, case):
    name, mesh, flux, bcp, u = case
    ctx = _ctx(mesh, M.PDE_EULER, flux, bc_prim=bcp)
    nb = ctx.Np * ctx.nv
    ud = torch.from_numpy(u).to(dev)
    blocks = torch.zeros(mesh.E * nb * nb, dtype=torch.float64, device=dev)
    ctx.assemble_diag_blocks(ud.data_ptr(), None, blocks.data_ptr())
    torch.cuda.synchronize()
    g = blocks.cpu().numpy().reshape(mesh.E, nb, nb)
    orc = O.Oracle(mesh, M.PDE_EULER, flux, bc_prim=bcp)
    ref = orc.diag_blocks(u)
    for e in range(mesh.E):
        diff = np.abs(g[e] - ref[e]).max()
        assert diff <= 4e-13 * max(1.0, np.abs(ref[e]).max()), (e, diff)
    # with the CFL mass term: A_e - s_e M_e (x) I
    s = np.linspace(0.5, 2.0, mesh.E)
    sd = torch.from_numpy(s).to(dev)
    ctx.assemble_diag_blocks(ud.data_ptr(), sd.data_ptr(), blocks.data_ptr())
    torch.cuda.synchronize()
    g2 = blocks.cpu().numpy().reshape(mesh.E, nb, nb)
    for e in (0, mesh.E - 1):
        Mm = np.kron(np.eye(ctx.nv), orc.mass(e))
        assert np.abs(g2[e] - (ref[e] - s[e] * Mm)).max() <= 1e-12 * np.abs(ref[e]).max()
    ctx.close()


@pytest.mark.parametrize("case", JAC, ids=[c[0] for c in JAC])
def test_jacobian_vector_matches_complex_step(dev, case):
    name, mesh, flux, bcp, u = case
    ctx = _ctx(mesh, M.PDE_EULER, flux, bc_prim=bcp)
    v = np.random.default_rng(9).standard_normal(u.size)
    ud, vd = torch.from_numpy(u).to(dev), torch.from_numpy(v).to(dev)
    jv = torch.empty_like(ud)
    ctx.jacobian_vector(ud.data_ptr(), vd.data_ptr(), jv.data_ptr())
    torch.cuda.synchronize()
    h = 1e-30
    _, im = O.Oracle(mesh, M.PDE_EULER, flux, bc_prim=bcp).rhs_complex(u, h * v)
    assert _rel(jv.cpu().numpy(), im / h) < 1e-12
    ctx.close()


def _neighbours(mesh):
    """(e, lf) -> neighbour element, from the face table rows (eL, fL, eR, fR, ord, bc)."""
    nb = {}
    for eL, fL, eR, fR, _, _ in mesh.faces:
        if eR >= 0:
            nb[(int(eL), int(fL))] = int(eR)
            nb[(int(eR), int(fR))] = int(eL)
    return nb


@pytest.mark.parametrize("case", JAC, ids=[c[0] for c in JAC])
def test_face_blocks_match_complex_step(dev, case):
    """dR_e/du_n for the face neighbours n against the oracle's complex-step block (the reference's
    lhs_lr / lhs_rl), and the whole block-sparse Jacobian applied to v against the exact J v."""
    name, mesh, flux, bcp, u = case
    ctx = _ctx(mesh, M.PDE_EULER, flux, bc_prim=bcp)
    nb, nf = ctx.Np * ctx.nv, 2 * mesh.d
    ud = torch.from_numpy(u).to(dev)
    fb = torch.zeros(mesh.E * nf * nb * nb, dtype=torch.float64, device=dev)
    db = torch.zeros(mesh.E * nb * nb, dtype=torch.float64, device=dev)
    ctx.assemble_face_blocks(ud.data_ptr(), fb.data_ptr())
    ctx.assemble_diag_blocks(ud.data_ptr(), None, db.data_ptr())
    v = np.random.default_rng(5).standard_normal(u.size)
    vd = torch.from_numpy(v).to(dev)
    jv = torch.empty_like(ud)
    ctx.jacobian_vector(ud.data_ptr(), vd.data_ptr(), jv.data_ptr())
    torch.cuda.synchronize()
    F = fb.cpu().numpy().reshape(mesh.E, nf, nb, nb)
    Dg = db.cpu().numpy().reshape(mesh.E, nb, nb)
    nbr = _neighbours(mesh)
    orc = O.Oracle(mesh, M.PDE_EULER, flux, bc_prim=bcp)
    for e in range(min(mesh.E, 3)):
        for n in sorted({nbr[(e, lf)] for lf in range(nf) if (e, lf) in nbr} - {e}):
            got = sum(F[e, lf] for lf in range(nf) if nbr.get((e, lf)) == n)
            ref = orc.jacobian_block(u, e, n)
            assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), (e, n)
    for e in range(mesh.E):                       # boundary faces carry no block
        for lf in range(nf):
            if (e, lf) not in nbr:
                assert not F[e, lf].any()
    V = v.reshape(mesh.E, nb)
    y = np.einsum("eij,ej->ei", Dg, V)
    for (e, lf), n in nbr.items():
        y[e] += F[e, lf] @ V[n]
    assert _rel(y.ravel(), jv.cpu().numpy()) < 1e-12
    ctx.close()


@pytest.mark.parametrize("case", JAC, ids=[c[0] for c in JAC])
def test_newton_operator_fuses_the_mass_term(dev, case):
    """dgx_newton_operator (J v - s_e M_e v in one dual-number launch, the GMRES operator) equals
    dgx_jacobian_vector followed by dgx_mass_sub."""
    name, mesh, flux, bcp, u = case
    ctx = _ctx(mesh, M.PDE_EULER, flux, bc_prim=bcp)
    rng = np.random.default_rng(19)
    v = rng.standard_normal(u.size)
    sc = rng.uniform(0.5, 30.0, mesh.E)
    ud, vd, sd = (torch.from_numpy(a).to(dev) for a in (u, v, sc))
    a = torch.empty_like(ud)
    b = torch.empty_like(ud)
    ctx.newton_operator(ud.data_ptr(), sd.data_ptr(), vd.data_ptr(), a.data_ptr())
    ctx.jacobian_vector(ud.data_ptr(), vd.data_ptr(), b.data_ptr())
    ctx.mass_sub(vd.data_ptr(), sd.data_ptr(), b.data_ptr())
    torch.cuda.synchronize()
    assert _rel(a.cpu().numpy(), b.cpu().numpy()) < 1e-13
    ctx.close()


def _bsr_apply(rp, ci, vals, v, nb):
    V = v.reshape(-1, nb)
    out = np.zeros_like(V)
    for e in range(len(rp) - 1):
        for k in range(rp[e], rp[e + 1]):
            out[e] += vals[k] @ V[ci[k]]
    return out.ravel()


BSR_CASES = JAC + [("periodic_2x1_lf_p2", M.structured_tp_mesh(2, 2, (2, 1), (0, 0), (1, 1), (True, True)),
                    M.FLUX_LF, None, None),
                   ("hex_periodic_2x2x2_roe_p2", M.structured_tp_mesh(3, 2, (2, 2, 2), (0, 0, 0), (1, 1, 1),
                                                                    (True, True, True)), M.FLUX_ROE, None, None)]


@pytest.mark.parametrize("case", BSR_CASES, ids=[c[0] for c in BSR_CASES])
def test_jacobian_bsr_is_the_newton_operator(dev, case):
    """dgx_assemble_jacobian_bsr in the pattern of dgx_jacobian_bsr_pattern (the reference's AIJ sparsity: each
    element's own block and its face neighbours', columns ascending) applied to v equals J v - s_e M_e v
    (dgx_newton_operator) and, without the mass term, J v; on a periodic mesh two elements wide and one high a
    neighbour reached through two faces, and the element itself, get the summed blocks."""
    name, mesh, flux, bcp, u = case
    if u is None:
        u = M.pack_state(M.prim_to_cons(M.smooth_state_3d(M.nodes_xyz(mesh)), mesh.d))
    ctx = _ctx(mesh, M.PDE_EULER, flux, bc_prim=bcp)
    nb = ctx.Np * ctx.nv
    rp, ci = ctx.jacobian_bsr_pattern()
    nbr = _neighbours(mesh)
    for e in range(mesh.E):
        want = sorted({e} | {n for (ee, lf), n in nbr.items() if ee == e})
        assert list(ci[rp[e]:rp[e + 1]]) == want
    ud = torch.from_numpy(u).to(dev)
    sc = np.random.default_rng(3).uniform(0.5, 5.0, mesh.E)
    sd = torch.from_numpy(sc).to(dev)
    vals = torch.full((len(ci) * nb * nb,), float("nan"), dtype=torch.float64, device=dev)
    v = np.random.default_rng(4).standard_normal(u.size)
    vd = torch.from_numpy(v).to(dev)
    ref = torch.empty_like(ud)
    for mass in (sd, None):
        ctx.assemble_jacobian_bsr(ud.data_ptr(), mass.data_ptr() if mass is not None else None, vals.data_ptr())
        if mass is not None:
            ctx.newton_operator(ud.data_ptr(), sd.data_ptr(), vd.data_ptr(), ref.data_ptr())
        else:
            ctx.jacobian_vector(ud.data_ptr(), vd.data_ptr(), ref.data_ptr())
        torch.cuda.synchronize()
        got = _bsr_apply(rp, ci, vals.cpu().numpy().reshape(-1, nb, nb), v, nb)
        assert _rel(got, ref.cpu().numpy()) < 1e-12, name
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


@pytest.mark.parametrize("case", JAC, ids=[c[0] for c in JAC])
def test_preconditioner_apply_matches_numpy(dev, case):
    """dgx_precond_apply: block-Jacobi is D^-1 v, and the multicolour block SGS sweep pair equals the
    same sweeps run in numpy over the assembled diagonal (with -M/dt) and face blocks; the SGS result
    is a better approximation of A^-1 v than block-Jacobi."""
    from dpgsolver_amd import _lib

    name, mesh, flux, bcp, u = case
    ctx = _ctx(mesh, M.PDE_EULER, flux, bc_prim=bcp)
    nb, nf = ctx.Np * ctx.nv, 2 * mesh.d
    ud = torch.from_numpy(u).to(dev)
    inv_dt = torch.empty(mesh.E, dtype=torch.float64, device=dev)
    ctx.cfl_dt(ud.data_ptr(), 5.0, inv_dt.data_ptr())
    fb = torch.zeros(mesh.E * nf * nb * nb, dtype=torch.float64, device=dev)
    db = torch.zeros(mesh.E * nb * nb, dtype=torch.float64, device=dev)
    ctx.assemble_face_blocks(ud.data_ptr(), fb.data_ptr())
    ctx.assemble_diag_blocks(ud.data_ptr(), inv_dt.data_ptr(), db.data_ptr())
    v = np.random.default_rng(23).standard_normal(u.size)
    vd = torch.from_numpy(v).to(dev)
    zb, zs = torch.empty_like(ud), torch.empty_like(ud)
    ctx.precond_apply(_lib.PC_BLOCK_JACOBI, ud.data_ptr(), inv_dt.data_ptr(), vd.data_ptr(), zb.data_ptr())
    ctx.precond_apply(_lib.PC_BLOCK_SGS, ud.data_ptr(), inv_dt.data_ptr(), vd.data_ptr(), zs.data_ptr())
    torch.cuda.synchronize()
    F = fb.cpu().numpy().reshape(mesh.E, nf, nb, nb)
    D = db.cpu().numpy().reshape(mesh.E, nb, nb)
    V = v.reshape(mesh.E, nb)
    nbr = _neighbours(mesh)
    Dinv_v = np.linalg.solve(D, V[..., None])[..., 0]
    assert _rel(zb.cpu().numpy(), Dinv_v.ravel()) < 1e-12
    col = _greedy_colours(mesh, nbr)
    colours = sorted(set(col))
    z = np.zeros_like(V)
    for sweep in (colours, colours[::-1]):       # forward, then backward
        for k in sweep:
            zold = z.copy()            # one colour at a time: its elements see the other colours only
            for e in (e for e in range(mesh.E) if col[e] == k):
                r = V[e].copy()
                for lf in range(nf):
                    if (e, lf) in nbr:
                        r -= F[e, lf] @ zold[nbr[(e, lf)]]
                z[e] = np.linalg.solve(D[e], r)
    got = zs.cpu().numpy()
    assert _rel(got, z.ravel()) < 1e-11
    # the same sweeps over FP32 copies of the blocks (products in FP64): FP32 storage rounding only
    z32 = torch.empty_like(ud)
    ctx.precond_apply(_lib.PC_BLOCK_SGS32, ud.data_ptr(), inv_dt.data_ptr(), vd.data_ptr(), z32.data_ptr())
    torch.cuda.synchronize()
    e32 = _rel(z32.cpu().numpy(), z.ravel())
    print(f"{name}: FP32-stored SGS vs FP64 sweeps {e32:.2e}")
    assert 0 < e32 < 1e-5
    A = np.zeros((mesh.E * nb, mesh.E * nb))
    for e in range(mesh.E):
        A[e * nb:(e + 1) * nb, e * nb:(e + 1) * nb] += D[e]
    for (e, lf), n in nbr.items():
        A[e * nb:(e + 1) * nb, n * nb:(n + 1) * nb] += F[e, lf]
    x = np.linalg.solve(A, v)
    assert np.linalg.norm(got - x) < np.linalg.norm(zb.cpu().numpy() - x)
    ctx.close()


def test_block_lu_solve_matches_numpy(dev):
    from dpgsolver_amd import _lib

    rng = np.random.default_rng(3)
    ctx = _lib.Context(2, 3, M.PDE_EULER, M.FLUX_ROE)
    # nb = 135 (HEX p = 2) and 143 take more than 64 KB of LDS per block: gfx950 workgroups hold up to 160 KB
    for E, nb in ((37, 64), (19, 135), (5, 143)):
        # eigenvalues within 1 of 4 at every size (circular law), so cond(A) stays O(1) and 1e-12 means the solve;
        # two columns swapped so that the first pivot is not on the diagonal
        A = rng.standard_normal((E, nb, nb)) / np.sqrt(nb) + 4 * np.eye(nb)
        A[:, :, [0, 5]] = A[:, :, [5, 0]]
        x = rng.standard_normal((E, nb))
        Ad = torch.from_numpy(A.copy()).to(dev)
        piv = torch.zeros(E * nb, dtype=torch.int32, device=dev)
        xd = torch.from_numpy(x.copy()).to(dev)
        ctx.block_lu_factor(Ad.data_ptr(), piv.data_ptr(), E, nb)
        ctx.block_lu_solve(Ad.data_ptr(), piv.data_ptr(), xd.data_ptr(), E, nb)
        torch.cuda.synchronize()
        ref = np.linalg.solve(A, x[..., None])[..., 0]
        assert _rel(xd.cpu().numpy(), ref) < 1e-12, nb
    # nb = 144: nb^2 doubles exceed the LDS; refused on the host before any launch (the buffers are 5 x 143^2)
    with pytest.raises(_lib.DgxError):
        ctx.block_lu_factor(Ad.data_ptr(), piv.data_ptr(), 1, 144)
    # the explicit (transposed) inverse used by the block-Jacobi and block SGS preconditioners, with pivoting
    # exercised; nb = 135 and 142 take more than 64 KB of LDS per block
    for E, nb in ((37, 64), (9, 135), (4, 142)):
        A = rng.standard_normal((E, nb, nb)) / np.sqrt(nb) + 4 * np.eye(nb)
        x = rng.standard_normal((E, nb))
        A2 = A.copy()
        A2[:, [0, 5]] = A2[:, [5, 0]]          # zero-free but pivot-demanding leading columns
        A2[:, 0, 0] = 1e-8
        Ai = torch.from_numpy(A2.copy()).to(dev)
        ctx.block_inverse(Ai.data_ptr(), E, nb)
        y = torch.empty((E, nb), dtype=torch.float64, device=dev)
        x2 = torch.from_numpy(x.copy()).to(dev)
        ctx.block_inv_apply(Ai.data_ptr(), x2.data_ptr(), y.data_ptr(), E, nb)
        torch.cuda.synchronize()
        inv = np.transpose(Ai.cpu().numpy().reshape(E, nb, nb), (0, 2, 1))
        np.testing.assert_allclose(inv, np.linalg.inv(A2), rtol=0, atol=1e-12 * np.abs(np.linalg.inv(A2)).max())
        ref2 = np.linalg.solve(A2, x[..., None])[..., 0]
        assert _rel(y.cpu().numpy(), ref2) < 1e-12, nb
    with pytest.raises(_lib.DgxError):   # refused on the host before any launch (the buffer is 4 x 142^2)
        ctx.block_inverse(Ai.data_ptr(), 1, 143)
    ctx.close()


def test_cfl_dt_closed_form_on_uniform_state(dev):
    mesh = M.channel_mesh(6, 4, 3)
    prim = M.freestream_prim(2)
    u = M.uniform_state(mesh, prim)
    ctx = _ctx(mesh, M.PDE_EULER, M.FLUX_ROE, bc_prim=prim)
    ud = torch.from_numpy(u).to(dev)
    inv = torch.empty(mesh.E, dtype=torch.float64, device=dev)
    ctx.cfl_dt(ud.data_ptr(), 7.0, inv.data_ptr())
    torch.cuda.synchronize()
    ws = abs(prim[1]) + np.sqrt(1.4 * prim[3] / prim[0])
    N = mesh.p + 1
    corners = [0, N - 1, N * (N - 1), N * N - 1]
    hmin = []
    for e in range(mesh.E):
        P = mesh.xyz[e][:, corners].T
        edges = [(0, 1), (2, 3), (0, 2), (1, 3)]
        hmin.append(min(np.linalg.norm(P[a] - P[b]) for a, b in edges))
    np.testing.assert_allclose(inv.cpu().numpy(), ws / (7.0 * np.array(hmin)), rtol=1e-13)
    ctx.close()


# ----------------------------------------------------------------------------- steady solve
def test_steady_channel_newton_gmres_block_jacobi(dev):
    from dpgsolver_amd import _lib

    mesh = M.channel_mesh(12, 6, 2)
    prim = M.freestream_prim(2)
    u = M.uniform_state(mesh, prim)
    ctx = _ctx(mesh, M.PDE_EULER, M.FLUX_ROE, bc_prim=prim)
    ud = torch.from_numpy(u).to(dev)
    st = ctx.solve_implicit(ud.data_ptr(), cfl_initial=10.0, exit_tol=1e-10, max_newton=80,
                            precond=_lib.PC_BLOCK_JACOBI)
    print("newton", st["newton_steps"], "gmres", st["gmres_iters"], "max_rhs", st["max_rhs"], st["seconds"])
    assert st["converged"] == 1 and st["max_rhs"] < 1e-10
    r = O.Oracle(mesh, M.PDE_EULER, M.FLUX_ROE, bc_prim=prim).rhs(ud.cpu().numpy())
    assert np.abs(r).max() < 1e-9          # the oracle agrees that this is the steady state
    ctx.close()


def test_steady_channel_block_sgs_fewer_iterations(dev):
    """The multicolour block SGS preconditioner reaches the same steady state as block-Jacobi with fewer
    GMRES iterations (it carries the face coupling that block-Jacobi drops); so does its variant with the
    blocks stored in FP32."""
    from dpgsolver_amd import _lib

    mesh = M.channel_mesh(12, 6, 2)
    prim = M.freestream_prim(2)
    ctx = _ctx(mesh, M.PDE_EULER, M.FLUX_ROE, bc_prim=prim)
    out = {}
    for pc in (_lib.PC_BLOCK_JACOBI, _lib.PC_BLOCK_SGS, _lib.PC_BLOCK_SGS32):
        ud = torch.from_numpy(M.uniform_state(mesh, prim)).to(dev)
        st = ctx.solve_implicit(ud.data_ptr(), cfl_initial=10.0, exit_tol=1e-10, max_newton=80, precond=pc)
        print("pc", pc, "newton", st["newton_steps"], "gmres", st["gmres_iters"], st["seconds"])
        assert st["converged"] == 1 and st["max_rhs"] < 1e-10
        out[pc] = (st, ud.cpu().numpy())
    assert out[_lib.PC_BLOCK_SGS][0]["gmres_iters"] < out[_lib.PC_BLOCK_JACOBI][0]["gmres_iters"]
    a, b = out[_lib.PC_BLOCK_SGS][1], out[_lib.PC_BLOCK_JACOBI][1]
    assert _rel(a, b) < 1e-7
    # FP32-stored factors: the same steady state, about as few iterations as the FP64 sweeps
    assert _rel(out[_lib.PC_BLOCK_SGS32][1], b) < 1e-7
    assert out[_lib.PC_BLOCK_SGS32][0]["gmres_iters"] <= 1.1 * out[_lib.PC_BLOCK_SGS][0]["gmres_iters"] + 5
    ctx.close()


def test_steady_channel_fp32_inner_operator(dev):
    """The GMRES operator in FP32 dual numbers (jvp_fp32: the inner defect-correction product over the FP32
    geometry), with the FP64 Newton residual deciding convergence: the same steady state as the FP64 operator,
    the same Newton steps, and GMRES iterations within 5%."""
    from dpgsolver_amd import _lib

    mesh = M.channel_mesh(12, 6, 2)
    prim = M.freestream_prim(2)
    ctx = _ctx(mesh, M.PDE_EULER, M.FLUX_ROE, bc_prim=prim)
    out = {}
    for fp32 in (False, True):
        ud = torch.from_numpy(M.uniform_state(mesh, prim)).to(dev)
        st = ctx.solve_implicit(ud.data_ptr(), cfl_initial=10.0, exit_tol=1e-10, max_newton=80,
                                precond=_lib.PC_BLOCK_JACOBI, jvp_fp32=fp32)
        print("jvp_fp32", fp32, "newton", st["newton_steps"], "gmres", st["gmres_iters"], st["seconds"])
        assert st["converged"] == 1 and st["max_rhs"] < 1e-10
        out[fp32] = (st, ud.cpu().numpy())
    (s64, u64), (s32, u32) = out[False], out[True]
    d = _rel(u32, u64)
    print(f"steady states rel diff {d:.2e}")
    assert d < 1e-9
    assert s32["newton_steps"] == s64["newton_steps"]
    assert abs(s32["gmres_iters"] - s64["gmres_iters"]) <= 0.05 * s64["gmres_iters"] + 4
    ctx.close()


def test_steady_solve_on_a_nonblocking_stream_matches_default_stream(dev):
    """The solver reads its norms and flags on the context's stream: a non-blocking caller stream (which does
    not order with the null stream) gives the same solve as the default stream."""
    from dpgsolver_amd import _lib

    mesh = M.channel_mesh(12, 6, 2)
    prim = M.freestream_prim(2)
    out = []
    for stream in (torch.cuda.default_stream(dev), torch.cuda.Stream(dev)):
        with torch.cuda.stream(stream):
            ctx = _lib.Context(2, 2, M.PDE_EULER, M.FLUX_ROE, bc_prim=prim)
            ctx.set_stream(stream.cuda_stream)
            ctx.upload_mesh(mesh.xyz, mesh.faces)
            ud = torch.from_numpy(M.uniform_state(mesh, prim)).to(dev)
            st = ctx.solve_implicit(ud.data_ptr(), cfl_initial=10.0, exit_tol=1e-10, max_newton=80,
                                    precond=_lib.PC_BLOCK_JACOBI)
            stream.synchronize()
            out.append((st, ud.cpu().numpy()))
            ctx.close()
    (s0, u0), (s1, u1) = out
    assert s0["converged"] == 1 and s1["converged"] == 1
    assert (s0["newton_steps"], s0["gmres_iters"]) == (s1["newton_steps"], s1["gmres_iters"])
    assert s0["history"] == s1["history"]
    np.testing.assert_array_equal(u0, u1)


def test_nonfinite_state_is_reported_not_solved(dev):
    """A NaN in the state makes the residual norm non-finite: DGX_ERR_NONFINITE, no Newton step taken
    (the reference would carry the NaN on; the library stops with a status code)."""
    from dpgsolver_amd import _lib

    mesh = M.channel_mesh(6, 4, 2)
    prim = M.freestream_prim(2)
    ctx = _ctx(mesh, M.PDE_EULER, M.FLUX_ROE, bc_prim=prim)
    u = M.uniform_state(mesh, prim)
    u[17] = np.nan
    ud = torch.from_numpy(u).to(dev)
    st = ctx.solve_implicit(ud.data_ptr(), max_newton=5, raise_on_error=False)
    assert st["status"] == -5 and st["newton_steps"] == 0 and st["converged"] == 0
    ctx.close()


def test_gmres_cap_is_counted_and_newton_continues(dev):
    """GMRES hitting its iteration cap is PETSc's DIVERGED_ITS: counted, and the Newton loop goes on
    (solve_implicit.c:668-672); the solve ends DGX_ERR_NOT_CONVERGED when the Newton steps run out."""
    from dpgsolver_amd import _lib

    mesh = M.channel_mesh(12, 6, 2)
    prim = M.freestream_prim(2)
    ctx = _ctx(mesh, M.PDE_EULER, M.FLUX_ROE, bc_prim=prim)
    u0, _ = _channel_state(mesh, amp=0.05)
    ud = torch.from_numpy(u0).to(dev)
    st = ctx.solve_implicit(ud.data_ptr(), cfl_initial=10.0, exit_tol=1e-14, max_newton=4, gmres_max_iter=2,
                            gmres_rtol=1e-12, precond=_lib.PC_BLOCK_JACOBI, raise_on_error=False)
    assert st["status"] == -6
    assert st["newton_steps"] == 4 and st["gmres_not_converged"] == 4
    assert st["gmres_per_step"] == [2, 2, 2, 2]
    assert st["history"][3] < st["history"][0]        # the capped steps still reduce the residual
    ctx.close()
