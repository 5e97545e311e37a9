"""GPU parity: the HIP path through the C-ABI against the CPU oracle (FP64 to 1e-12 relative).

Cases cover the configuration shapes (2-D advection p=2, 2-D Euler p=3 Roe, 3-D Euler p=4 LF on
a curved HEX mesh, 2-D channel with Riemann/slip-wall boundaries) at sizes the oracle finishes in
seconds, plus the edge cases of the face coupling: single-element periodic meshes (an element is
its own neighbour), rotated element-local numbering (every face orientation), and 1e-6 perturbed
states (non-equilibrium residual).
"""
import numpy as np
import pytest

import pyoracle as O
from dpgsolver_amd import mesh as M

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL64 = 1e-12


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


def _curved2d(n, p, amp=0.04, rotate=None):
    def curve(pts):
        out = pts.copy()
        x, y = pts[..., 0], pts[..., 1]
        out[..., 0] = x + amp * np.sin(2 * np.pi * y)
        out[..., 1] = y + amp * np.sin(2 * np.pi * x)
        return out
    return M.structured_tp_mesh(2, p, (n, n), (0.0, 0.0), (1.0, 1.0), (True, True), curve=curve, rotate=rotate)


def _cases():
    rng = np.random.default_rng(0)
    out = []
    m = M.config1_mesh(8, 2)
    out.append(("adv2d_p2", m, M.PDE_ADVECTION, M.FLUX_UPWIND, (1.0, 0.5, 0.0), None, M.config1_state(m)))
    m = M.config2_mesh(16, 3)
    out.append(("euler2d_p3_roe_vortex", m, M.PDE_EULER, M.FLUX_ROE, (0, 0, 0), None, M.config2_state(m)))
    m = _curved2d(5, 3, rotate=rng.integers(0, 4, 25))
    u = M.pack_state(M.prim_to_cons(M.smooth_state_3d(M.nodes_xyz(m)), 2))
    out.append(("euler2d_p3_lf_curved_rotated", m, M.PDE_EULER, M.FLUX_LF, (0, 0, 0), None, u))
    m = M.config3_mesh(3, 4)
    out.append(("euler3d_p4_lf_curved", m, M.PDE_EULER, M.FLUX_LF, (0, 0, 0), None, M.config3_state(m)))
    m = M.structured_tp_mesh(3, 2, (3, 3, 3), (0,) * 3, (1,) * 3, (True,) * 3, curve=M.curve3d(0.05),
                             rotate=rng.integers(0, 8, 27))
    u = M.pack_state(M.prim_to_cons(M.smooth_state_3d(M.nodes_xyz(m)), 3))
    out.append(("euler3d_p2_roe_rotated", m, M.PDE_EULER, M.FLUX_ROE, (0, 0, 0), None, u))
    m = M.channel_mesh(8, 6, 3)
    prim = M.freestream_prim(2)
    x = M.nodes_xyz(m)
    u = M.pack_state(M.prim_to_cons(prim[None, None, :] + 0.02 * np.sin(3 * x[..., :1]) * np.array([1, 1, 0.5, 1]),
                                    2))
    out.append(("euler2d_p3_channel_bc", m, M.PDE_EULER, M.FLUX_ROE, (0, 0, 0), prim, u))
    m = M.structured_tp_mesh(2, 3, (1, 1), (0, 0), (1, 1), (True, True))
    u = M.pack_state(M.prim_to_cons(M.smooth_state_3d(M.nodes_xyz(m)), 2))
    out.append(("euler2d_single_element", m, M.PDE_EULER, M.FLUX_ROE, (0, 0, 0), None, u))
    m = M.structured_tp_mesh(3, 4, (2, 1, 2), (0,) * 3, (1,) * 3, (True,) * 3, curve=M.curve3d(0.03))
    u = M.pack_state(M.prim_to_cons(M.smooth_state_3d(M.nodes_xyz(m)), 3))
    out.append(("euler3d_p4_tiny_self_neighbour", m, M.PDE_EULER, M.FLUX_LF, (0, 0, 0), None, u))
    for p in (1, 2, 4, 5):
        m = M.config2_mesh(4, p)
        out.append((f"euler2d_p{p}_roe", m, M.PDE_EULER, M.FLUX_ROE, (0, 0, 0), None, M.config2_state(m)))
    for p in (1, 3):
        m = M.config3_mesh(2, p)
        out.append((f"euler3d_p{p}_roe", m, M.PDE_EULER, M.FLUX_ROE, (0, 0, 0), None, M.config3_state(m)))
    return out


CASES = _cases()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_geometry_matches_oracle(dev, case):
    name, mesh, pde, flux, b, bcp, u = case
    ctx = _ctx(mesh, pde, flux, b, bcp)
    met, dj, nJ = ctx.geometry()
    om, od, on = O.Oracle(mesh, pde, flux, b, bcp).geometry()
    assert _rel(met, om) < TOL64
    assert _rel(dj, od) < TOL64
    assert _rel(nJ, on) < TOL64
    ctx.close()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_fp64_residual_matches_oracle(dev, case):
    name, mesh, pde, flux, b, bcp, u = case
    print(f"case {name}: E={mesh.E}", flush=True)
    ctx = _ctx(mesh, pde, flux, b, bcp)
    ud = torch.from_numpy(u).to(dev)
    rd = torch.full_like(ud, float("nan"))
    ctx.compute_rhs(ud.data_ptr(), rd.data_ptr())
    torch.cuda.synchronize()
    ref = O.Oracle(mesh, pde, flux, b, bcp).rhs(u)
    assert np.isfinite(rd.cpu().numpy()).all()
    assert _rel(rd.cpu().numpy(), ref) < TOL64, _rel(rd.cpu().numpy(), ref)
    ctx.close()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_volume_and_face_terms_separately_match_oracle(dev, case):
    """dgx_compute_volume_rhs / dgx_compute_face_rhs (compute_volume_rlhs_dg / compute_face_rlhs_dg) against the
    oracle's or_volume_rhs_d / or_face_rhs_d; both accumulate into r like the reference, and the two calls on a
    zeroed r give the full residual."""
    name, mesh, pde, flux, b, bcp, u = case
    print(f"case {name}: E={mesh.E}", flush=True)
    ctx = _ctx(mesh, pde, flux, b, bcp)
    orc = O.Oracle(mesh, pde, flux, b, bcp)
    ud = torch.from_numpy(u).to(dev)
    vol_ref, face_ref = orc.rhs(u, "volume"), orc.rhs(u, "face")
    scale = max(np.abs(vol_ref).max(), np.abs(face_ref).max())
    rv = torch.zeros_like(ud)
    ctx.compute_volume_rhs(ud.data_ptr(), rv.data_ptr())
    rf = torch.zeros_like(ud)
    ctx.compute_face_rhs(ud.data_ptr(), rf.data_ptr())
    offset = torch.full_like(ud, 0.25)                   # accumulation: r += term
    ctx.compute_face_rhs(ud.data_ptr(), offset.data_ptr())
    both = torch.zeros_like(ud)
    ctx.compute_volume_rhs(ud.data_ptr(), both.data_ptr())
    ctx.compute_face_rhs(ud.data_ptr(), both.data_ptr())
    full = torch.empty_like(ud)
    ctx.compute_rhs(ud.data_ptr(), full.data_ptr())
    torch.cuda.synchronize()
    assert np.abs(rv.cpu().numpy() - vol_ref).max() < TOL64 * scale
    assert np.abs(rf.cpu().numpy() - face_ref).max() < TOL64 * scale
    # r += f: only the rounding of 0.25 + f (half an ulp of |0.25 + f|) separates it from 0.25 + rf
    assert np.abs((offset - 0.25 - rf).cpu().numpy()).max() <= 2.0 ** -52 * (0.25 + rf.abs().max().item())
    assert np.abs((both - full).cpu().numpy()).max() < TOL64 * scale
    ctx.close()


@pytest.mark.parametrize("case", CASES[:6], ids=[c[0] for c in CASES[:6]])
def test_explicit_rhs_with_mass_inverse(dev, case):
    name, mesh, pde, flux, b, bcp, u = case
    ctx = _ctx(mesh, pde, flux, b, bcp)
    orc = O.Oracle(mesh, pde, flux, b, bcp)
    ref, ref_max = orc.rhs_explicit(u)
    ud = torch.from_numpy(u).to(dev)
    rd = torch.empty_like(ud)
    mx = ctx.compute_rhs_explicit(ud.data_ptr(), rd.data_ptr())
    assert _rel(rd.cpu().numpy(), ref) < 1e-11
    assert abs(mx - ref_max) <= 1e-11 * ref_max
    # standalone M^-1 on the plain residual
    r2 = torch.empty_like(ud)
    ctx.compute_rhs(ud.data_ptr(), r2.data_ptr())
    ctx.apply_minv(r2.data_ptr())
    torch.cuda.synchronize()
    assert _rel(r2.cpu().numpy(), ref) < 1e-11
    ctx.close()


def test_fp32_residual_close_to_fp64(dev):
    mesh = M.config3_mesh(4, 4)
    u = M.config3_state(mesh)
    ctx = _ctx(mesh, M.PDE_EULER, M.FLUX_LF)
    ud = torch.from_numpy(u).to(dev)
    r64 = torch.empty_like(ud)
    ctx.compute_rhs(ud.data_ptr(), r64.data_ptr())
    u32 = torch.empty(ud.numel(), dtype=torch.float32, device=dev)
    ctx.convert_f64_f32(ud.data_ptr(), u32.data_ptr(), ud.numel())
    r32 = torch.empty_like(u32)
    ctx.compute_rhs_f32(u32.data_ptr(), r32.data_ptr())
    torch.cuda.synchronize()
    err = _rel(r32.double().cpu().numpy(), r64.cpu().numpy())
    # the floor any FP32-state kernel has: the FP64 residual of the state rounded to FP32 (R is ~60x smaller than
    # the volume and face terms that cancel in it); the kernel's own FP32 arithmetic adds rounding of the same
    # size (measured 2.8x the floor on the 64^3 workload)
    rr = torch.empty_like(ud)
    ur = u32.double()
    ctx.compute_rhs(ur.data_ptr(), rr.data_ptr())
    torch.cuda.synchronize()
    floor = _rel(rr.cpu().numpy(), r64.cpu().numpy())
    print(f"FP32 vs FP64 relative error (inf norm): {err:.3e}, input rounding alone {floor:.3e}")
    assert floor > 0.0
    assert err < 6.0 * floor, (err, floor)
    ctx.close()


EULER = [c for c in CASES if c[2] == M.PDE_EULER]


@pytest.mark.parametrize("case", EULER, ids=[c[0] for c in EULER])
def test_mixed_residual_matches_oracle(dev, case):
    """FP32 contractions in per-element perturbation form, FP64 state and pointwise fluxes."""
    name, mesh, pde, flux, b, bcp, u = case
    ctx = _ctx(mesh, pde, flux, b, bcp)
    ud = torch.from_numpy(u).to(dev)
    r = torch.empty(ud.numel(), dtype=torch.float32, device=dev)
    ctx.compute_rhs_mixed(ud.data_ptr(), r.data_ptr())
    torch.cuda.synchronize()
    ref = O.Oracle(mesh, pde, flux, b, bcp).rhs(u)
    err = _rel(r.double().cpu().numpy(), ref)
    print(f"{name}: mixed vs oracle {err:.3e}")
    assert err < 2e-5
    ctx.close()


def test_mixed_residual_error_on_curved_hex(dev):
    """At 16^3 curved HEX p=4 the perturbation form keeps the FP32 error below 1e-6 (the pure FP32
    path, whose state is rounded to FP32, does not)."""
    mesh = M.config3_mesh(16, 4)
    u = M.config3_state(mesh)
    ctx = _ctx(mesh, M.PDE_EULER, M.FLUX_LF)
    ud = torch.from_numpy(u).to(dev)
    r64 = torch.empty_like(ud)
    ctx.compute_rhs(ud.data_ptr(), r64.data_ptr())
    rm = torch.empty(ud.numel(), dtype=torch.float32, device=dev)
    ctx.compute_rhs_mixed(ud.data_ptr(), rm.data_ptr())
    u32 = torch.empty(ud.numel(), dtype=torch.float32, device=dev)
    ctx.convert_f64_f32(ud.data_ptr(), u32.data_ptr(), ud.numel())
    r32 = torch.empty_like(u32)
    ctx.compute_rhs_f32(u32.data_ptr(), r32.data_ptr())
    torch.cuda.synchronize()
    ref = r64.cpu().numpy()
    em, e32 = _rel(rm.double().cpu().numpy(), ref), _rel(r32.double().cpu().numpy(), ref)
    print(f"16^3 HEX p=4: mixed {em:.3e}, fp32 {e32:.3e}")
    assert em < 1e-6 and em < 0.5 * e32
    ctx.close()


@pytest.mark.parametrize("name", ["straight_2d_quad_periodic", "parametric_2d_quad_periodic_reflected"])
def test_residual_on_ingested_gmsh_mesh(dev, name):
    """A reference test mesh read by dgx_mesh_read_gmsh drives the GPU residual: parity with the oracle on
    the same face table, and free-stream preservation on the fully periodic one (the other has walls that
    the straight-sided geometry does not align with the flow)."""
    import os
    from dpgsolver_amd import _lib

    g = _lib.read_gmsh(os.path.join(os.path.dirname(__file__), "golden", "mesh", name + ".msh"), 3)
    mesh = M.Mesh(2, 3, g["xyz"], g["faces"].astype(np.int32))
    prim = M.freestream_prim(2)
    x = M.nodes_xyz(mesh)
    u = M.pack_state(M.prim_to_cons(prim[None, None, :] + 0.05 * np.sin(2 * x[..., :1] + x[..., 1:2])
                                    * np.array([1, 1, 0.5, 1]), 2))
    for flux in (M.FLUX_ROE, M.FLUX_LF):
        ctx = _ctx(mesh, M.PDE_EULER, flux, bc_prim=prim)
        ud = torch.from_numpy(u).to(dev)
        r = torch.empty_like(ud)
        ctx.compute_rhs(ud.data_ptr(), r.data_ptr())
        fs = torch.from_numpy(M.uniform_state(mesh, prim)).to(dev)
        rf = torch.empty_like(fs)
        ctx.compute_rhs(fs.data_ptr(), rf.data_ptr())
        torch.cuda.synchronize()
        ref = O.Oracle(mesh, M.PDE_EULER, flux, bc_prim=prim).rhs(u)
        assert _rel(r.cpu().numpy(), ref) < TOL64
        if name.startswith("straight"):          # fully periodic: the free stream is a steady state
            assert rf.abs().max().item() < 1e-13
        ctx.close()


@pytest.mark.parametrize("periodic", [True, False], ids=["periodic", "riemann"])
def test_residual_on_ingested_hex_gmsh_mesh(dev, tmp_path, periodic):
    """A HEX gmsh mesh (elements listing their vertices after random cube rotations, so every face orientation
    code occurs) read by dgx_mesh_read_gmsh drives the GPU residual: parity with the oracle on the same face table."""
    from dpgsolver_amd import _lib
    from test_mesh_ingest import _write_hex_gmsh

    path = str(tmp_path / "hex.msh")
    _write_hex_gmsh(path, 3, periodic, seed=5)
    g = _lib.read_gmsh(path, 3)
    mesh = M.Mesh(3, 3, g["xyz"], g["faces"].astype(np.int32))
    prim = M.freestream_prim(3)
    u = M.pack_state(M.prim_to_cons(M.smooth_state_3d(M.nodes_xyz(mesh)), 3))
    for flux in (M.FLUX_ROE, M.FLUX_LF):
        ctx = _ctx(mesh, M.PDE_EULER, flux, bc_prim=prim)
        ud = torch.from_numpy(u).to(dev)
        r = torch.full_like(ud, float("nan"))
        ctx.compute_rhs(ud.data_ptr(), r.data_ptr())
        torch.cuda.synchronize()
        ref = O.Oracle(mesh, M.PDE_EULER, flux, bc_prim=prim).rhs(u)
        assert _rel(r.cpu().numpy(), ref) < TOL64
        ctx.close()


def test_free_stream_preserved_on_curved_hex_gpu(dev):
    mesh = M.config3_mesh(6, 4)
    prim = np.array([1.0, 0.3, -0.2, 0.1, 1.0])
    u = M.uniform_state(mesh, prim)
    ctx = _ctx(mesh, M.PDE_EULER, M.FLUX_LF)
    ud = torch.from_numpy(u).to(dev)
    rd = torch.empty_like(ud)
    ctx.compute_rhs(ud.data_ptr(), rd.data_ptr())
    torch.cuda.synchronize()
    assert rd.abs().max().item() < 2e1 * 1e-15 * 10
    ctx.close()
