"""GPU parity of the variable-order 2-D Euler residual (dgx_upload_mesh_hp) against the oracle (hp_oracle.py).

* FP64: the face kernel between orders plus one residual launch per order group equals the oracle's
  per-volume / per-face residual to 1e-12 relative: random orders 1..5 (seeded), Roe and Lax-Friedrichs,
  straight and curved periodic meshes with rotated elements, and a channel with Riemann and slip-wall faces.
* A mesh of one order through the variable-order path equals the uniform context bit for bit.
* The explicit path (M^-1 R, its max-norm, SSP-RK3 stages) equals the oracle's, step by step.
* FP32 agrees with FP64 to FP32 rounding, the mixed form (FP64 state and fluxes, FP32 contractions on the
  perturbation, the faces between orders in FP64) at least twice as closely; the residual sums to zero on a
  periodic mesh.
* Navier-Stokes (BR2): the groups' gradient kernels, the face kernel between orders (projected face products
  and lifted fluxes), the groups' residual kernels equal the oracle's rhs_ns to 1e-12; FP32 to FP32 rounding.
* Unsorted orders, orders above the context's p and viscous meshes with boundary faces are refused; entry
  points that do not serve variable-order meshes return DGX_ERR_UNSUPPORTED.
"""
import numpy as np
import pytest

import hp_oracle as H
from dpgsolver_amd import mesh as M

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def _curve(pts):
    out = pts.copy()
    x, y = pts[..., 0], pts[..., 1]
    out[..., 0] = x + 0.03 * np.sin(2 * np.pi * y)
    out[..., 1] = y + 0.03 * np.sin(2 * np.pi * x)
    return out


def _state_fn(seed):
    ph = np.random.default_rng(seed).uniform(0, 2 * np.pi, 4)

    def f(x):
        X, Y = x[:, 0], x[:, 1]
        prim = np.stack([1 + 0.1 * np.sin(2 * np.pi * X + ph[0]) * np.cos(2 * np.pi * Y),
                         0.3 + 0.2 * np.sin(2 * np.pi * Y + ph[1]), -0.1 + 0.2 * np.cos(2 * np.pi * X + ph[2]),
                         1 + 0.1 * np.sin(2 * np.pi * (X + Y) + ph[3])], -1)
        return M.prim_to_cons(prim, 2)
    return f


def _ctx(mesh, flux, pmax=5, bc_prim=None):
    from dpgsolver_amd import _lib

    kw = {} if bc_prim is None else {"bc_prim": bc_prim}
    ctx = _lib.Context(2, pmax, M.PDE_EULER, flux, **kw)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    ctx.upload_mesh_hp(mesh)
    return ctx


CASES = [(n, flux, curved, rot) for n in ((5, 4), (6, 6)) for flux in (M.FLUX_ROE, M.FLUX_LF)
         for curved in (False, True) for rot in (False, True)]


@pytest.mark.parametrize("n,flux,curved,rot", CASES,
                         ids=[f"{n[0]}x{n[1]}_{'roe' if f == M.FLUX_ROE else 'lf'}{'_curved' if c else ''}"
                              f"{'_rot' if r else ''}" for n, f, c, r in CASES])
def test_hp_residual_matches_oracle(dev, n, flux, curved, rot):
    E = n[0] * n[1]
    pe = M.config4_hp_orders(n, seed=E + 7 * curved + 3 * rot)
    rotate = np.random.default_rng(E).integers(0, 4, E) if rot else None
    mesh = M.structured_hp_mesh(pe, n, (0, 0), (1, 1), (True, True), curve=_curve if curved else None,
                                rotate=rotate)
    u = M.hp_state(mesh, _state_fn(E))
    ref = H.HpOracle(mesh, M.PDE_EULER, flux).rhs(u)
    ctx = _ctx(mesh, flux)
    assert ctx.ndof == u.size
    ud = torch.from_numpy(u).to(dev)
    r = torch.full_like(ud, np.nan)
    ctx.compute_rhs(ud.data_ptr(), r.data_ptr())
    torch.cuda.synchronize()
    got = r.cpu().numpy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"{n} orders {np.bincount(pe, minlength=6)[1:]} flux={flux} curved={curved} rot={rot}: {err:.2e}")
    assert err < 1e-12
    off = mesh.offsets(4)
    tot = sum(got[off[e]:off[e + 1]].reshape(4, -1).sum(1) for e in range(mesh.E))
    assert np.abs(tot).max() < 1e-13 * np.abs(got).sum()
    # FP32: FP32 state, geometry and operators
    u32 = ud.float()
    r32 = torch.full_like(u32, np.nan)
    ctx.compute_rhs_f32(u32.data_ptr(), r32.data_ptr())
    torch.cuda.synchronize()
    e32 = ((r32.double() - r).abs().max() / r.abs().max()).item()
    assert e32 < 2e-5, e32
    # mixed: FP64 state, FP32 contractions on u - c_e, FP64 fluxes (the faces between orders in FP64)
    rmx = torch.full_like(u32, np.nan)
    ctx.compute_rhs_mixed(ud.data_ptr(), rmx.data_ptr())
    torch.cuda.synchronize()
    emx = ((rmx.double() - r).abs().max() / r.abs().max()).item()
    print(f"   FP32 {e32:.2e}  mixed {emx:.2e}")
    assert emx < 0.5 * e32, (emx, e32)
    ctx.close()


def test_hp_channel_with_boundary_faces(dev):
    prim = M.freestream_prim(2)
    n = (6, 4)
    pe = M.config4_hp_orders(n, seed=11)

    def curve(pts):
        out = pts.copy()
        x, eta = pts[..., 0], pts[..., 1] / 0.8
        b = 0.0625 * np.exp(-25.0 * x * x)
        out[..., 1] = b + eta * (0.8 - b)
        return out

    mesh = M.structured_hp_mesh(pe, n, (-1.5, 0.0), (1.5, 0.8), (False, False), curve=curve,
                                bc_low=(M.BC_RIEMANN, M.BC_SLIPWALL), bc_high=(M.BC_RIEMANN, M.BC_SLIPWALL))
    u = M.hp_state(mesh, lambda x: M.prim_to_cons(np.tile(prim, (x.shape[0], 1)), 2)
                   * (1 + 0.02 * np.sin(3 * x[:, :1]) * np.cos(2 * x[:, 1:])))
    for flux in (M.FLUX_ROE, M.FLUX_LF):
        ref = H.HpOracle(mesh, M.PDE_EULER, flux, bc_prim=prim).rhs(u)
        ctx = _ctx(mesh, flux, bc_prim=prim)
        ud = torch.from_numpy(u).to(dev)
        r = torch.full_like(ud, np.nan)
        ctx.compute_rhs(ud.data_ptr(), r.data_ptr())
        torch.cuda.synchronize()
        err = np.abs(r.cpu().numpy() - ref).max() / np.abs(ref).max()
        assert err < 1e-12, err
        ctx.close()


def test_one_order_through_hp_equals_uniform_context(dev):
    from dpgsolver_amd import _lib

    n = (5, 4)
    for p in (2, 4):
        hm = M.structured_hp_mesh(np.full(20, p), n, (0, 0), (1, 1), (True, True), curve=_curve)
        um = M.structured_tp_mesh(2, p, n, (0, 0), (1, 1), (True, True), curve=_curve)
        u = torch.from_numpy(M.hp_state(hm, _state_fn(p))).to(dev)
        ctx_h = _ctx(hm, M.FLUX_ROE)
        ctx_u = _lib.Context(2, p, M.PDE_EULER, M.FLUX_ROE)
        ctx_u.set_stream(torch.cuda.current_stream().cuda_stream)
        ctx_u.upload_mesh(um.xyz, um.faces)
        rh, ru = torch.empty_like(u), torch.empty_like(u)
        ctx_h.compute_rhs(u.data_ptr(), rh.data_ptr())
        ctx_u.compute_rhs(u.data_ptr(), ru.data_ptr())
        torch.cuda.synchronize()
        assert torch.equal(rh, ru)
        ctx_h.close()
        ctx_u.close()


def test_hp_explicit_path_matches_oracle(dev):
    n = (5, 5)
    pe = M.config4_hp_orders(n, seed=3)
    mesh = M.structured_hp_mesh(pe, n, (0, 0), (1, 1), (True, True), rotate=np.random.default_rng(2).integers(0, 4, 25))
    u = M.hp_state(mesh, _state_fn(4))
    orc = H.HpOracle(mesh, M.PDE_EULER, M.FLUX_ROE)
    ctx = _ctx(mesh, M.FLUX_ROE)
    ud = torch.from_numpy(u).to(dev)
    r = torch.empty_like(ud)
    mx = ctx.compute_rhs_explicit(ud.data_ptr(), r.data_ptr())
    torch.cuda.synchronize()
    ref = orc.apply_minv(orc.rhs(u))
    assert np.abs(r.cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max()
    assert abs(mx - np.abs(ref).max()) <= 1e-12 * np.abs(ref).max()
    # SSP-RK3 (Gottlieb 2001 eq. 4.2) without the limiter, against the same stages on the oracle
    dt = 2e-3
    uref = u.copy()
    for _ in range(2):
        ctx.ssp_rk3_step(ud.data_ptr(), dt, positivity=False)
        f = lambda w: orc.apply_minv(orc.rhs(w))
        s1 = uref + dt * f(uref)
        s2 = 0.75 * uref + 0.25 * (s1 + dt * f(s1))
        uref = uref / 3.0 + 2.0 / 3.0 * (s2 + dt * f(s2))
    torch.cuda.synchronize()
    assert np.abs(ud.cpu().numpy() - uref).max() <= 1e-12 * np.abs(uref).max()
    # with the limiter on a smooth state nothing is limited
    u2 = ud.clone()
    ctx.ssp_rk3_step(u2.data_ptr(), dt, positivity=True)
    assert torch.isfinite(u2).all()
    ctx.close()


def test_hp_refusals(dev):
    from dpgsolver_amd import _lib

    n = (4, 4)
    pe = M.config4_hp_orders(n)
    mesh = M.structured_hp_mesh(pe, n, (0, 0), (1, 1), (True, True))
    bad = M.HpMesh(p=mesh.p[::-1].copy(), xyz=mesh.xyz[::-1], faces=mesh.faces)
    ctx = _lib.Context(2, 5, M.PDE_EULER, M.FLUX_ROE)
    with pytest.raises(_lib.DgxError, match="ARG"):
        ctx.upload_mesh_hp(bad)                     # orders not sorted
    ctx.close()
    ctx = _lib.Context(2, 3, M.PDE_EULER, M.FLUX_ROE)
    with pytest.raises(_lib.DgxError, match="ARG"):
        ctx.upload_mesh_hp(mesh)                    # an order above the context's p
    ctx.close()
    ch = M.structured_hp_mesh(pe, n, (0, 0), (1, 1), (False, False), bc_low=(M.BC_RIEMANN, M.BC_SLIPWALL),
                              bc_high=(M.BC_RIEMANN, M.BC_SLIPWALL))
    ctx = _lib.Context(2, 5, M.PDE_NAVIER_STOKES, M.FLUX_ROE, mu=0.01)
    with pytest.raises(_lib.DgxError, match="UNSUPPORTED"):
        ctx.upload_mesh_hp(ch)                      # the slip wall is an Euler condition
    ctx.close()
    ctx = _ctx(mesh, M.FLUX_ROE)
    u = torch.from_numpy(M.hp_state(mesh, _state_fn(0))).to(dev)
    r = torch.empty_like(u)
    with pytest.raises(_lib.DgxError, match="UNSUPPORTED"):
        ctx.jacobian_vector(u.data_ptr(), u.data_ptr(), r.data_ptr())
    ctx.close()


MU, PR = 0.02, 0.72
NS_CASES = [(n, flux, curved, rot) for n in ((5, 4), (6, 6)) for flux in (M.FLUX_ROE, M.FLUX_LF)
            for curved in (False, True) for rot in (False, True)]


def _ns_ctx(mesh, flux):
    from dpgsolver_amd import _lib

    ctx = _lib.Context(2, 5, M.PDE_NAVIER_STOKES, flux, mu=MU, Pr=PR)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    ctx.upload_mesh_hp(mesh)
    return ctx


@pytest.mark.parametrize("n,flux,curved,rot", NS_CASES,
                         ids=[f"{n[0]}x{n[1]}_{'roe' if f == M.FLUX_ROE else 'lf'}{'_curved' if c else ''}"
                              f"{'_rot' if r else ''}" for n, f, c, r in NS_CASES])
def test_hp_ns_residual_matches_oracle(dev, n, flux, curved, rot):
    E = n[0] * n[1]
    pe = M.config4_hp_orders(n, seed=E + 5 * curved + 2 * rot)
    rotate = np.random.default_rng(E + 1).integers(0, 4, E) if rot else None
    mesh = M.structured_hp_mesh(pe, n, (0, 0), (1, 1), (True, True), curve=_curve if curved else None,
                                rotate=rotate)
    u = M.hp_state(mesh, _state_fn(E + 1))
    orc = H.HpOracle(mesh, M.PDE_EULER, flux, viscous=(MU, PR))
    ref = orc.rhs_ns(u)
    assert np.abs(ref - orc.rhs(u)).max() > 1e-3 * np.abs(ref).max()   # the viscous part is not negligible
    ctx = _ns_ctx(mesh, flux)
    ud = torch.from_numpy(u).to(dev)
    r = torch.full_like(ud, np.nan)
    ctx.compute_rhs(ud.data_ptr(), r.data_ptr())
    torch.cuda.synchronize()
    got = r.cpu().numpy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"{n} orders {np.bincount(pe, minlength=6)[1:]} flux={flux} curved={curved} rot={rot}: {err:.2e}")
    assert err < 1e-12
    off = mesh.offsets(4)
    tot = sum(got[off[e]:off[e + 1]].reshape(4, -1).sum(1) for e in range(mesh.E))
    assert np.abs(tot).max() < 1e-13 * np.abs(got).sum()
    u32 = ud.float()
    r32 = torch.full_like(u32, np.nan)
    ctx.compute_rhs_f32(u32.data_ptr(), r32.data_ptr())
    torch.cuda.synchronize()
    e32 = ((r32.double() - r).abs().max() / r.abs().max()).item()
    assert e32 < 2e-4, e32          # FP32 state rounding, amplified by the second derivatives
    rmx = torch.full_like(u32, np.nan)
    ctx.compute_rhs_mixed(ud.data_ptr(), rmx.data_ptr())
    torch.cuda.synchronize()
    emx = ((rmx.double() - r).abs().max() / r.abs().max()).item()
    print(f"   FP32 {e32:.2e}  mixed {emx:.2e}")
    assert emx < 0.5 * e32, (emx, e32)
    ctx.close()


def test_hp_ns_explicit_path_and_uniform_equivalence(dev):
    from dpgsolver_amd import _lib

    n = (5, 5)
    pe = M.config4_hp_orders(n, seed=9)
    mesh = M.structured_hp_mesh(pe, n, (0, 0), (1, 1), (True, True), curve=_curve)
    u = M.hp_state(mesh, _state_fn(8))
    orc = H.HpOracle(mesh, M.PDE_EULER, M.FLUX_ROE, viscous=(MU, PR))
    ctx = _ns_ctx(mesh, M.FLUX_ROE)
    ud = torch.from_numpy(u).to(dev)
    r = torch.empty_like(ud)
    mx = ctx.compute_rhs_explicit(ud.data_ptr(), r.data_ptr())
    torch.cuda.synchronize()
    ref = orc.apply_minv(orc.rhs_ns(u))
    assert np.abs(r.cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max()
    assert abs(mx - np.abs(ref).max()) <= 1e-12 * np.abs(ref).max()
    # one SSP-RK3 step (Gottlieb 2001 eq. 4.2, no limiter) against the same stages on the oracle
    dt = 1e-3
    f = lambda w: orc.apply_minv(orc.rhs_ns(w))
    s1 = u + dt * f(u)
    s2 = 0.75 * u + 0.25 * (s1 + dt * f(s1))
    uref = u / 3.0 + 2.0 / 3.0 * (s2 + dt * f(s2))
    ctx.ssp_rk3_step(ud.data_ptr(), dt, positivity=False)
    torch.cuda.synchronize()
    assert np.abs(ud.cpu().numpy() - uref).max() <= 1e-12 * np.abs(uref).max()
    ctx.close()
    # one order through the variable-order path equals the uniform Navier-Stokes context bit for bit
    hm = M.structured_hp_mesh(np.full(20, 3), (5, 4), (0, 0), (1, 1), (True, True), curve=_curve)
    um = M.structured_tp_mesh(2, 3, (5, 4), (0, 0), (1, 1), (True, True), curve=_curve)
    uu = torch.from_numpy(M.hp_state(hm, _state_fn(3))).to(dev)
    ch = _ns_ctx(hm, M.FLUX_ROE)
    cu = _lib.Context(2, 3, M.PDE_NAVIER_STOKES, M.FLUX_ROE, mu=MU, Pr=PR)
    cu.set_stream(torch.cuda.current_stream().cuda_stream)
    cu.upload_mesh(um.xyz, um.faces)
    rh, ru = torch.empty_like(uu), torch.empty_like(uu)
    ch.compute_rhs(uu.data_ptr(), rh.data_ptr())
    cu.compute_rhs(uu.data_ptr(), ru.data_ptr())
    torch.cuda.synchronize()
    assert torch.equal(rh, ru)
    ch.close()
    cu.close()


def test_hp_graph_replays_equal_eager_calls(dev):
    from dpgsolver_amd import _lib

    n = (6, 5)
    pe = M.config4_hp_orders(n, seed=21)
    mesh = M.structured_hp_mesh(pe, n, (0, 0), (1, 1), (True, True), curve=_curve)
    for pde in (M.PDE_EULER, M.PDE_NAVIER_STOKES):
        ctx = _lib.Context(2, 5, pde, M.FLUX_LF, mu=MU, Pr=PR)
        ctx.set_stream(torch.cuda.current_stream().cuda_stream)
        ctx.upload_mesh_hp(mesh)
        u64 = torch.from_numpy(M.hp_state(mesh, _state_fn(5))).to(dev)
        u32 = u64.float()
        outs = {_lib.GRAPH_RHS_FP64: torch.empty_like(u64), _lib.GRAPH_RHS_FP32: torch.empty_like(u32),
                _lib.GRAPH_RHS_MIXED: torch.empty_like(u32)}
        ins = {_lib.GRAPH_RHS_FP64: u64, _lib.GRAPH_RHS_FP32: u32, _lib.GRAPH_RHS_MIXED: u64}
        eager = {_lib.GRAPH_RHS_FP64: ctx.compute_rhs, _lib.GRAPH_RHS_FP32: ctx.compute_rhs_f32,
                 _lib.GRAPH_RHS_MIXED: ctx.compute_rhs_mixed}
        graphs = {k: ctx.graph_rhs(k, ins[k].data_ptr(), outs[k].data_ptr()) for k in outs}
        for rep in range(3):
            for k, g in graphs.items():
                ref = torch.empty_like(outs[k])
                eager[k](ins[k].data_ptr(), ref.data_ptr())
                torch.cuda.synchronize()
                outs[k].fill_(np.nan)
                g.launch(1)
                g.sync()
                assert torch.equal(outs[k], ref), (pde, k, rep)
            u64.mul_(1.0 + 1e-3)
            u32.copy_(u64)
        for g in graphs.values():
            g.close()
        ctx.close()


def test_hp_residual_on_a_caller_stream(dev):
    """A variable-order residual spreads its launches over the context's side streams and joins them back to the
    caller's stream: on a caller stream other than the default one, work enqueued on that stream right after
    the call (a copy of the residual) sees the whole residual, eagerly and through a graph, bitwise equal to
    the same call on the default stream."""
    from dpgsolver_amd import _lib

    n = (6, 5)
    pe = M.config4_hp_orders(n, seed=23)
    mesh = M.structured_hp_mesh(pe, n, (0, 0), (1, 1), (True, True), curve=_curve)
    side = torch.cuda.Stream(device=dev)
    for pde in (M.PDE_EULER, M.PDE_NAVIER_STOKES):
        ctx = _lib.Context(2, 5, pde, M.FLUX_ROE, mu=MU, Pr=PR)
        ctx.set_stream(torch.cuda.current_stream().cuda_stream)
        ctx.upload_mesh_hp(mesh)
        u64 = torch.from_numpy(M.hp_state(mesh, _state_fn(7))).to(dev)
        u32 = u64.float()
        calls = [(ctx.compute_rhs, u64, torch.float64), (ctx.compute_rhs_f32, u32, torch.float32),
                 (ctx.compute_rhs_mixed, u64, torch.float32)]
        refs = []
        for fn, ui, dt in calls:
            r = torch.empty(u64.shape, dtype=dt, device=dev)
            fn(ui.data_ptr(), r.data_ptr())
            refs.append(r)
        torch.cuda.synchronize()
        side.wait_stream(torch.cuda.current_stream())
        ctx.set_stream(side.cuda_stream)
        with torch.cuda.stream(side):
            for rep in range(2):
                for (fn, ui, dt), ref in zip(calls, refs):
                    r = torch.full(u64.shape, float("nan"), dtype=dt, device=dev)
                    fn(ui.data_ptr(), r.data_ptr())
                    seen = r.clone()   # enqueued on the caller stream right after the residual
                    side.synchronize()
                    assert torch.equal(seen, ref), (pde, fn.__name__, rep)
            out = torch.full(u64.shape, float("nan"), dtype=torch.float64, device=dev)
            g = ctx.graph_rhs(_lib.GRAPH_RHS_FP64, u64.data_ptr(), out.data_ptr())
            g.launch(1)
            seen = out.clone()
            side.synchronize()
            assert torch.equal(seen, refs[0]), (pde, "graph")
            g.close()
        ctx.close()


def test_hp_failed_upload_leaves_no_mesh(dev):
    """An upload that fails in a group (a viscous mesh with boundary faces) leaves the context without a mesh:
    the residual then reports DGX_ERR_NO_MESH instead of running on a half-built variable-order state."""
    from dpgsolver_amd import _lib

    n = (4, 4)
    pe = M.config4_hp_orders(n)
    ch = M.structured_hp_mesh(pe, n, (0, 0), (1, 1), (False, False), bc_low=(M.BC_RIEMANN, M.BC_SLIPWALL),
                              bc_high=(M.BC_RIEMANN, M.BC_SLIPWALL))
    ctx = _lib.Context(2, 5, M.PDE_NAVIER_STOKES, M.FLUX_ROE, mu=MU, Pr=PR)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    with pytest.raises(_lib.DgxError, match="UNSUPPORTED"):
        ctx.upload_mesh_hp(ch)
    u = torch.from_numpy(M.hp_state(ch, _state_fn(1))).to(dev)
    r = torch.empty_like(u)
    with pytest.raises(_lib.DgxError, match="NO_MESH"):
        ctx.compute_rhs(u.data_ptr(), r.data_ptr())
    ctx.close()
