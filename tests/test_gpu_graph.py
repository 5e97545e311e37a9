"""Launch machinery of the residual kernels.

* HIP graphs of the explicit step and of the residual (dgx_graph_*): every replay equals the eager call
  bit for bit.  Each test replays a graph several times with eager work and synchronisations in between,
  and checks every replay.
* Persistent launch grids (dgx_set_rhs_grid) and the launch-shape autotuner (dgx_autotune_rhs): the
  residual does not depend on the grid beyond rounding, and the autotuner keeps the results."""
import numpy as np
import pytest

from dpgsolver_amd import mesh as M

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def _ctx(mesh, pde, flux, b_adv=(0.0, 0.0, 0.0)):
    from dpgsolver_amd import _lib

    ctx = _lib.Context(mesh.d, mesh.p, pde, flux, b_adv=b_adv)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    ctx.upload_mesh(mesh.xyz, mesh.faces)
    return ctx


@pytest.mark.parametrize("case", ["adv", "euler_vortex"])
def test_rk3_graph_replays_equal_eager_steps(dev, case):
    if case == "adv":
        mesh = M.config1_mesh(8, 2)
        u = M.config1_state(mesh)
        pde, flux, b, dt, pos = M.PDE_ADVECTION, M.FLUX_UPWIND, (1.0, 0.5, 0.0), 0.2 * 0.25 / (5 * 1.0), False
    else:
        mesh = M.config2_mesh(8, 3)
        u = M.config2_state(mesh)
        pde, flux, b, dt, pos = M.PDE_EULER, M.FLUX_ROE, (0.0, 0.0, 0.0), 0.01, True
    ctx = _ctx(mesh, pde, flux, b)
    ue = torch.from_numpy(u).to(dev)
    ug = ue.clone()
    other = torch.empty_like(ue)
    g = ctx.graph_rk3(ug.data_ptr(), dt, positivity=pos)
    for rep in range(4):
        mx_e = ctx.ssp_rk3_step(ue.data_ptr(), dt, positivity=pos)
        ctx.compute_rhs(ue.data_ptr(), other.data_ptr())      # eager work between the replays
        torch.cuda.synchronize()
        g.launch(1)
        mx_g = g.sync()
        assert mx_g == mx_e, rep
        assert torch.equal(ue, ug), rep
    # several replays in one launch call equal as many eager steps
    for _ in range(3):
        ctx.ssp_rk3_step(ue.data_ptr(), dt, positivity=pos)
    g.launch(3)
    g.sync()
    assert torch.equal(ue, ug)
    g.close()
    ctx.close()


def test_residual_graphs_equal_eager_calls(dev):
    from dpgsolver_amd import _lib

    mesh = M.config3_mesh(4, 3)
    u = M.config3_state(mesh)
    ctx = _ctx(mesh, M.PDE_EULER, M.FLUX_LF)
    u64 = torch.from_numpy(u).to(dev)
    u32 = u64.float()
    outs = {k: (torch.empty_like(u64) if k == _lib.GRAPH_RHS_FP64 else torch.empty_like(u32))
            for k in (_lib.GRAPH_RHS_FP64, _lib.GRAPH_RHS_FP32, _lib.GRAPH_RHS_MIXED)}
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
            assert torch.equal(outs[k], ref), (k, rep)
        u64.mul_(1.0 + 1e-3)           # the graphs read the buffers they were recorded with
        u32.copy_(u64)
    for g in graphs.values():
        g.close()
    ctx.close()


@pytest.mark.parametrize("grid", [8, 64, 256])
def test_persistent_grids_give_the_same_residuals(dev, grid):
    """The persistent kernel is a separate compilation of the same code: the compiler may pair other
    multiplies and adds into FMAs, so it agrees with the one-block kernel to rounding (FP64 1e-13, FP32
    1e-6 relative to max|R|), not bit for bit; every output is written (the outputs start as NaN)."""
    from dpgsolver_amd import _lib

    mesh = M.config3_mesh(4, 3)
    u = M.config3_state(mesh)
    ctx = _ctx(mesh, M.PDE_EULER, M.FLUX_LF)
    u64 = torch.from_numpy(u).to(dev)
    u32 = u64.float()
    calls = {_lib.GRAPH_RHS_FP64: (ctx.compute_rhs, u64, torch.float64, 1e-13),
             _lib.GRAPH_RHS_FP32: (ctx.compute_rhs_f32, u32, torch.float32, 1e-6),
             _lib.GRAPH_RHS_MIXED: (ctx.compute_rhs_mixed, u64, torch.float32, 1e-6)}
    for kind, (fn, uin, dt, tol) in calls.items():
        ref = torch.full((u.size,), np.nan, dtype=dt, device=dev)
        got = torch.full_like(ref, np.nan)
        ctx.set_rhs_grid(kind, 0)
        fn(uin.data_ptr(), ref.data_ptr())
        ctx.set_rhs_grid(kind, grid)
        fn(uin.data_ptr(), got.data_ptr())
        torch.cuda.synchronize()
        assert not torch.isnan(got).any() and not torch.isnan(ref).any(), kind
        err = ((got.double() - ref.double()).abs().max() / ref.double().abs().max()).item()
        assert err <= tol, (kind, err)
    # the explicit step's M^-1 R launch, with its max-norm reduction, on a 2-D Euler mesh
    mesh2 = M.config2_mesh(8, 3)
    ctx2 = _ctx(mesh2, M.PDE_EULER, M.FLUX_ROE)
    ua = torch.from_numpy(M.config2_state(mesh2)).to(dev)
    ub = ua.clone()
    for _ in range(3):
        ctx2.set_rhs_grid(_lib.TUNE_RHS_MINV, 0)
        ma = ctx2.ssp_rk3_step(ua.data_ptr(), 0.01)
        ctx2.set_rhs_grid(_lib.TUNE_RHS_MINV, grid)
        mb = ctx2.ssp_rk3_step(ub.data_ptr(), 0.01)
        assert abs(ma - mb) <= 1e-13 * ma
    assert ((ua - ub).abs().max() / ua.abs().max()).item() <= 1e-13
    ctx2.close()
    ctx.close()


def test_autotuner_picks_a_listed_grid_and_keeps_results(dev):
    from dpgsolver_amd import _lib

    mesh = M.config3_mesh(8, 4)
    u = M.config3_state(mesh)
    ctx = _ctx(mesh, M.PDE_EULER, M.FLUX_LF)
    u32 = torch.from_numpy(u).to(dev).float()
    ref = torch.empty_like(u32)
    ctx.compute_rhs_f32(u32.data_ptr(), ref.data_ptr())
    scratch = torch.empty_like(u32)
    grid, ms = ctx.autotune_rhs(_lib.GRAPH_RHS_FP32, u32.data_ptr(), scratch.data_ptr())
    ncu = torch.cuda.get_device_properties(dev).multi_processor_count
    assert grid == 0 or grid in [m * ncu for m in (1, 2, 3, 4, 6, 8, 12, 16)]
    assert ms > 0.0
    got = torch.full_like(u32, np.nan)
    ctx.compute_rhs_f32(u32.data_ptr(), got.data_ptr())
    torch.cuda.synchronize()
    assert not torch.isnan(got).any()
    assert ((got - ref).abs().max() / ref.abs().max()).item() <= 1e-6   # equal up to FMA pairing
    ctx.close()
