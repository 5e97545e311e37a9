"""GPU parity of the 2-D Navier-Stokes (BR2) residual against the oracle (or_compute_rhs_ns).

* The FP64 residual (k_ns_grad + k_rhs_ns) equals the oracle's dense-operator residual to 1e-12 relative on
  curved periodic meshes, Roe and Lax-Friedrichs, p = 1..5, at a state whose viscous part is a sizeable
  fraction of the residual (checked).
* The explicit path (M^-1 R and its max-norm, an SSP-RK3 step) uses the same kernels.
* The face coupling is conservative: the residual sums to zero per conserved variable on a periodic mesh.
* The FP32 residual (FP32 state, geometry and operators) agrees with the FP64 one to FP32 rounding, and the
  mixed residual (FP64 state, FP32 contractions on the perturbation from the element's first node) closer
  still (the gap widens with refinement: 4x at 4x4, see the bench's 448x448 leg).
* Boundary faces: the no-slip walls (adiabatic, diabatic, rotating) on the Taylor-Couette annulus and Riemann
  faces on a box equal the oracle to 1e-12 (FP64 residual, M^-1 R), p = 1..5, both fluxes; the FP32 and mixed
  forms stay within a multiple of the input-rounding floor.
* Slip walls (an Euler condition) and the learned preconditioner (trained on Euler blocks) are refused with
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


def _mesh(n, p, amp=0.03, rotate=None):
    def curve(pts):
        out = pts.copy()
        x, y = pts[..., 0], pts[..., 1]
        out[..., 0] = x + amp * np.sin(2 * np.pi * y)
        out[..., 1] = y + amp * np.sin(2 * np.pi * x)
        return out
    return M.structured_tp_mesh(2, p, (n, n), (0.0, 0.0), (1.0, 1.0), (True, True), curve=curve, rotate=rotate)


def _state(mesh, seed=0):
    x = M.nodes_xyz(mesh)
    ph = np.random.default_rng(seed).uniform(0, 2 * np.pi, 4)
    X, Y = x[..., 0], x[..., 1]
    rho = 1.0 + 0.1 * np.sin(2 * np.pi * X + ph[0]) * np.cos(2 * np.pi * Y)
    u = 0.3 + 0.2 * np.sin(2 * np.pi * Y + ph[1])
    v = -0.1 + 0.2 * np.cos(2 * np.pi * X + ph[2])
    p = 1.0 + 0.1 * np.sin(2 * np.pi * (X + Y) + ph[3])
    return M.pack_state(M.prim_to_cons(np.stack([rho, u, v, p], -1), 2))


def _ctx(mesh, flux, walls=None, bc_prim=None):
    from dpgsolver_amd import _lib

    ctx = _lib.Context(2, mesh.p, M.PDE_NAVIER_STOKES, flux, mu=MU, Pr=PR, walls=walls, bc_prim=bc_prim)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    ctx.upload_mesh(mesh.xyz, mesh.faces)
    return ctx


CASES = [(n, p, flux, rot) for (n, p) in ((6, 1), (5, 2), (4, 3), (3, 4), (3, 5))
         for flux in (M.FLUX_ROE, M.FLUX_LF) for rot in (False, True)]


@pytest.mark.parametrize("n,p,flux,rot", CASES,
                         ids=[f"n{n}_p{p}_{'roe' if f == M.FLUX_ROE else 'lf'}{'_rot' if r else ''}"
                              for n, p, f, r in CASES])
def test_ns_residual_matches_oracle(dev, n, p, flux, rot):
    rotate = np.random.default_rng(n * 10 + p).integers(0, 4, n * n) if rot else None
    mesh = _mesh(n, p, rotate=rotate)
    u = _state(mesh, seed=p)
    orc = O.Oracle(mesh, M.PDE_EULER, flux, viscous=(MU, PR))
    ref = orc.rhs_ns(u)
    ref_euler = O.Oracle(mesh, M.PDE_EULER, flux).rhs(u)
    assert np.abs(ref - ref_euler).max() > 1e-3 * np.abs(ref).max()   # the viscous part is not negligible
    ctx = _ctx(mesh, flux)
    ud = torch.from_numpy(u).to(dev)
    r = torch.empty_like(ud)
    ctx.compute_rhs(ud.data_ptr(), r.data_ptr())
    torch.cuda.synchronize()
    got = r.cpu().numpy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"n={n} p={p} flux={flux} rot={rot}: rel err {err:.2e}")
    assert err < 1e-12
    # conservation: each conserved variable's residual sums to zero over a periodic mesh
    tot = got.reshape(mesh.E, 4, -1).sum(axis=(0, 2))
    assert np.abs(tot).max() < 1e-12 * np.abs(got).sum()
    ctx.close()


def test_ns_explicit_path_matches_oracle(dev):
    mesh = _mesh(4, 3)
    u = _state(mesh, seed=5)
    orc = O.Oracle(mesh, M.PDE_EULER, M.FLUX_ROE, viscous=(MU, PR))
    ref = orc.apply_minv(orc.rhs_ns(u))
    ctx = _ctx(mesh, M.FLUX_ROE)
    ud = torch.from_numpy(u).to(dev)
    r = torch.empty_like(ud)
    mx = ctx.compute_rhs_explicit(ud.data_ptr(), r.data_ptr())
    torch.cuda.synchronize()
    assert np.abs(r.cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max()
    assert abs(mx - np.abs(ref).max()) <= 1e-12 * np.abs(ref).max()
    # SSP-RK3 steps with the limiter: stable, and the integral of every conserved variable is kept
    dt = 2e-3
    for _ in range(10):
        ctx.ssp_rk3_step(ud.data_ptr(), dt)
    torch.cuda.synchronize()
    u1 = ud.cpu().numpy()
    assert np.isfinite(u1).all()
    mass0 = sum(orc.mass(e).sum(axis=0) @ u.reshape(mesh.E, 4, -1)[e].T for e in range(mesh.E))
    mass1 = sum(orc.mass(e).sum(axis=0) @ u1.reshape(mesh.E, 4, -1)[e].T for e in range(mesh.E))
    assert np.abs(mass1 - mass0).max() < 1e-12 * np.abs(mass0).max(), (mass0, mass1)
    ctx.close()


def test_ns_refuses_boundary_faces_and_unbuilt_paths(dev):
    from dpgsolver_amd import _lib

    ch = M.channel_mesh(4, 3, 2)
    ctx = _lib.Context(2, 2, M.PDE_NAVIER_STOKES, M.FLUX_ROE, mu=MU, Pr=PR)
    with pytest.raises(_lib.DgxError, match="UNSUPPORTED"):
        ctx.upload_mesh(ch.xyz, ch.faces)
    ctx.close()
    mesh = _mesh(3, 2)
    ctx = _ctx(mesh, M.FLUX_ROE)
    u = torch.from_numpy(_state(mesh)).to(dev)
    with pytest.raises(_lib.DgxError, match="UNSUPPORTED"):
        ctx.solve_implicit(u.data_ptr(), precond=_lib.PC_LEARNED)
    ctx.close()


@pytest.mark.parametrize("p", [2, 3, 5])
def test_ns_fp32_residual_close_to_fp64(dev, p):
    mesh = _mesh(4, p, rotate=np.random.default_rng(p).integers(0, 4, 16))
    u = _state(mesh, seed=p)
    ctx = _ctx(mesh, M.FLUX_ROE)
    u64 = torch.from_numpy(u).to(dev)
    u32 = u64.float()
    r64 = torch.empty_like(u64)
    r32 = torch.empty_like(u32)
    ctx.compute_rhs(u64.data_ptr(), r64.data_ptr())
    ctx.compute_rhs_f32(u32.data_ptr(), r32.data_ptr())
    torch.cuda.synchronize()
    err = ((r32.double() - r64).abs().max() / r64.abs().max()).item()
    print(f"p={p}: FP32 vs FP64 rel err {err:.2e}")
    assert err < 2e-4          # FP32 rounding of O(1) volume and face terms that cancel to the residual
    # mixed form: FP64 state, FP32 contractions on u - c_e, FP64 pointwise fluxes
    rmx = torch.empty_like(u32)
    ctx.compute_rhs_mixed(u64.data_ptr(), rmx.data_ptr())
    torch.cuda.synchronize()
    errmx = ((rmx.double() - r64).abs().max() / r64.abs().max()).item()
    print(f"p={p}: mixed vs FP64 rel err {errmx:.2e}")
    assert errmx < 0.5 * err and errmx < 5e-6    # measured 7.3e-7 against 3.0e-6 (p=2, 4x4)
    tot = r32.double().reshape(mesh.E, 4, -1).sum(dim=(0, 2))
    assert tot.abs().max().item() < 1e-5 * r32.double().abs().sum().item()
    ctx.close()


# ---------------------------------------------------------------------------------------------- boundary faces
WALL_CASES = [(p, n, dT, flux) for (p, n) in ((1, 3), (2, 3), (3, 2), (4, 2), (5, 2)) for dT in (0.0, 1.0)
              for flux in (M.FLUX_ROE, M.FLUX_LF)]


def _wall_state(mesh, dTdr_o):
    """The Taylor-Couette state plus a smooth perturbation (every wall term active, R far from zero)."""
    u, _ = M.taylor_couette_state(mesh, MU, PR, dTdr_o=dTdr_o)
    x = M.nodes_xyz(mesh)
    pert = 0.02 * np.sin(3 * x[..., 0]) * np.cos(2 * x[..., 1])
    return u * (1.0 + M.pack_state(np.stack([pert, 2 * pert, -pert, 0.5 * pert], -1)))


@pytest.mark.parametrize("p,n,dTdr_o,flux", WALL_CASES,
                         ids=[f"p{p}_n{n}_{'diabatic' if dT else 'adiabatic'}_{'roe' if f == M.FLUX_ROE else 'lf'}"
                              for p, n, dT, f in WALL_CASES])
def test_ns_no_slip_walls_match_oracle(dev, p, n, dTdr_o, flux):
    bc_outer = M.BC_NOSLIP_DIABATIC if dTdr_o else M.BC_NOSLIP_ADIABATIC
    mesh = M.annulus_mesh(p, n, 4 * n, bc_outer=bc_outer)
    walls = M.taylor_couette_walls(MU, PR, dTdr_o=dTdr_o)
    u = _wall_state(mesh, dTdr_o)
    print(f"annulus p={p} {n}x{4 * n} outer={'diabatic' if dTdr_o else 'adiabatic'} flux={flux}", flush=True)
    orc = O.Oracle(mesh, M.PDE_EULER, flux, viscous=(MU, PR), walls=walls)
    ref = orc.rhs_ns(u)
    ctx = _ctx(mesh, flux, walls=walls)
    ud = torch.from_numpy(u).to(dev)
    r = torch.full_like(ud, float("nan"))
    ctx.compute_rhs(ud.data_ptr(), r.data_ptr())
    torch.cuda.synchronize()
    got = r.cpu().numpy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"  FP64 rel err {err:.2e}", flush=True)
    assert err < 1e-12
    # explicit path: M^-1 R and its max-norm
    rx = torch.full_like(ud, float("nan"))
    mx = ctx.compute_rhs_explicit(ud.data_ptr(), rx.data_ptr())
    torch.cuda.synchronize()
    refx = orc.apply_minv(ref)
    assert np.abs(rx.cpu().numpy() - refx).max() <= 1e-12 * np.abs(refx).max()
    assert abs(mx - np.abs(refx).max()) <= 1e-12 * np.abs(refx).max()
    # FP32 and mixed forms against FP64, bounded by the input-rounding floor
    u32 = ud.float()
    rr = torch.empty_like(ud)
    ctx.compute_rhs(u32.double().data_ptr(), rr.data_ptr())
    r32 = torch.full_like(u32, float("nan"))
    rmx = torch.full_like(u32, float("nan"))
    ctx.compute_rhs_f32(u32.data_ptr(), r32.data_ptr())
    ctx.compute_rhs_mixed(ud.data_ptr(), rmx.data_ptr())
    torch.cuda.synchronize()
    scale = r.abs().max().item()
    floor = (rr - r).abs().max().item() / scale
    e32 = (r32.double() - r).abs().max().item() / scale
    emx = (rmx.double() - r).abs().max().item() / scale
    print(f"  FP32 {e32:.2e}, mixed {emx:.2e}, input-rounding floor {floor:.2e}", flush=True)
    assert e32 < 8 * floor and emx < e32
    ctx.close()


def test_ns_riemann_box_matches_oracle(dev):
    prim = M.freestream_prim(2)
    curve = lambda pts: np.stack([pts[..., 0] + 0.03 * np.sin(2 * np.pi * pts[..., 1]),
                                  pts[..., 1] + 0.03 * np.sin(2 * np.pi * pts[..., 0])], -1)
    mesh = M.structured_tp_mesh(2, 3, (5, 4), (0.0, 0.0), (1.0, 0.8), (False, False), curve=curve,
                                bc_low=(M.BC_RIEMANN, M.BC_RIEMANN), bc_high=(M.BC_RIEMANN, M.BC_RIEMANN))
    x = M.nodes_xyz(mesh)
    pert = 0.02 * np.sin(3 * x[..., 0]) * np.cos(2 * x[..., 1])
    u = M.pack_state(M.prim_to_cons(prim[None, None, :] * (1.0 + np.stack([pert, pert, -pert, 0.5 * pert], -1)), 2))
    for flux in (M.FLUX_ROE, M.FLUX_LF):
        orc = O.Oracle(mesh, M.PDE_EULER, flux, bc_prim=prim, viscous=(MU, PR))
        ref = orc.rhs_ns(u)
        ctx = _ctx(mesh, flux, bc_prim=prim)
        ud = torch.from_numpy(u).to(dev)
        r = torch.full_like(ud, float("nan"))
        ctx.compute_rhs(ud.data_ptr(), r.data_ptr())
        torch.cuda.synchronize()
        err = np.abs(r.cpu().numpy() - ref).max() / np.abs(ref).max()
        print(f"Riemann box flux={flux}: rel err {err:.2e}", flush=True)
        assert err < 1e-12
        # a free stream stays a free stream
        uf = torch.from_numpy(M.uniform_state(mesh, prim)).to(dev)
        ctx.compute_rhs(uf.data_ptr(), r.data_ptr())
        torch.cuda.synchronize()
        assert r.abs().max().item() < 1e-13
        ctx.close()
