"""GPU checks of the learned block preconditioner.

* The MFMA (bf16 x bf16 -> fp32) forward pass against the oracle's bf16-emulating float64 forward.
* Loss and parameter gradient of the hand-written backward kernels against the oracle's analytic
  gradient (itself pinned by finite differences in test_mlp_oracle.py).
* Features and the analytic scale sigma against the oracle on a channel mesh.
* Training on the diagonal blocks the solver captures along its own Newton iterations lowers the loss,
  and the trained network, used inside GMRES, reaches the same 1e-10 FP64 steady state as the exact LU
  block-Jacobi solve with fewer GMRES iterations than the untrained network (whose output is close to
  the analytic scaled identity -sigma I).  Both structures are covered: diagonal + rank 8 and the
  Kronecker sum of rank 8.
"""
import numpy as np
import pytest

import mlp_oracle as MO
import pyoracle as O
from dpgsolver_amd import mesh as M

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def _mlp(nb, Np, seed=7, kind=0):
    from dpgsolver_amd import _lib

    return _lib.MLP(nb, Np, hidden=256, seed=seed, kind=kind)


KINDS = [0, 1]   # low rank (diag + rank 8), Kronecker sum (rank 8)


def _perturbed_params(m, rng):
    w = m.get_params().astype(np.float64)
    w += 0.05 * rng.standard_normal(w.size)
    return w.astype(np.float32)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("nb,Np,B", [(64, 16, 37), (36, 9, 50)])
def test_forward_matches_oracle(dev, nb, Np, B, kind):
    rng = np.random.default_rng(21)
    m = _mlp(nb, Np, kind=kind)
    w = _perturbed_params(m, rng)
    m.set_params(w)
    X = rng.standard_normal((B, m.nf)).astype(np.float32)
    Xd = torch.from_numpy(X).to(dev)
    out = torch.empty(B * m.nout, dtype=torch.float32, device=dev)
    m.predict(Xd.data_ptr(), B, out.data_ptr())
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(B, m.nout).astype(np.float64)
    _, _, ref = MO.forward(w, X, m.nf, m.hidden, m.nout, emulate_bf16=True)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("forward rel err vs bf16-emulating oracle", err)
    assert err < 3e-3
    m.close()


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("nb,Np,B", [(64, 16, 29), (36, 9, 40)])
def test_loss_gradient_matches_oracle(dev, nb, Np, B, kind):
    rng = np.random.default_rng(22)
    m = _mlp(nb, Np, kind=kind)
    w = _perturbed_params(m, rng)
    m.set_params(w)
    X = rng.standard_normal((B, m.nf)).astype(np.float32)
    A = -(rng.standard_normal((B, nb, nb)) * 0.3 + 2.0 * np.eye(nb))
    sigma = rng.uniform(0.3, 0.7, B).astype(np.float32)
    Xd, Ad, sd = (torch.from_numpy(a).to(dev) for a in (X, A, sigma))
    loss, g = m.loss_grad(Xd.data_ptr(), sd.data_ptr(), Ad.data_ptr(), B)
    Lref, gref = MO.loss_and_grad(w, X, sigma, A, nb, m.rank, m.hidden, emulate_bf16=True, kind=kind, Np=Np)
    L64, g64 = MO.loss_and_grad(w, X, sigma, A, nb, m.rank, m.hidden, emulate_bf16=False, kind=kind, Np=Np)
    gerr = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    print(f"loss {loss:.6e} ref {Lref:.6e}; grad rel err {gerr:.2e} (vs pure float64: "
          f"{np.linalg.norm(g - g64) / np.linalg.norm(g64):.2e})")
    assert abs(loss - Lref) <= 1e-4 * Lref
    assert gerr < 2e-3
    offs, _, n = MO.layout(m.nf, m.hidden, m.nout)
    for a, b in zip(offs, offs[1:] + [n]):        # every parameter tensor, not only the largest
        assert np.linalg.norm(g[a:b] - gref[a:b]) <= 5e-3 * np.linalg.norm(gref[a:b]) + 1e-12
    m.close()


def _channel(nx, ny, p=3):
    mesh = M.channel_mesh(nx, ny, p)
    prim = M.freestream_prim(2)
    return mesh, prim


def _ctx(mesh, prim):
    from dpgsolver_amd import _lib

    ctx = _lib.Context(mesh.d, mesh.p, M.PDE_EULER, M.FLUX_ROE, bc_prim=prim)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    ctx.upload_mesh(mesh.xyz, mesh.faces)
    return ctx


def _hmin(mesh):
    N = mesh.p + 1
    c = [0, N - 1, N * (N - 1), N * N - 1]
    P = mesh.xyz[:, :, c]                                    # (E, 2, 4)
    edges = [(0, 1), (2, 3), (0, 2), (1, 3)]
    return np.min([np.linalg.norm(P[:, :, a] - P[:, :, b], axis=1) for a, b in edges], axis=0)


def test_features_match_oracle(dev):
    mesh, prim = _channel(8, 4)
    ctx = _ctx(mesh, prim)
    m = _mlp(64, 16)
    x = M.nodes_xyz(mesh)
    pert = 0.05 * np.sin(3 * x[..., :1]) * np.cos(2 * x[..., 1:2]) * np.array([1, 1, 0.5, 1])
    u = M.pack_state(M.prim_to_cons(prim[None, None, :] + pert, 2))
    ud = torch.from_numpy(u).to(dev)
    inv = torch.empty(mesh.E, dtype=torch.float64, device=dev)
    ctx.cfl_dt(ud.data_ptr(), 50.0, inv.data_ptr())
    feat = torch.empty(mesh.E * m.nf, dtype=torch.float32, device=dev)
    sig = torch.empty(mesh.E, dtype=torch.float32, device=dev)
    m.features(ctx, ud.data_ptr(), inv.data_ptr(), feat.data_ptr(), sig.data_ptr())
    torch.cuda.synchronize()
    _, detJ, _ = O.Oracle(mesh, M.PDE_EULER, M.FLUX_ROE, bc_prim=prim).geometry()
    rho, vx, vy, p = prim[:4]
    vel = np.hypot(vx, vy)
    uref = (rho, rho * vel, p / 0.4 + 0.5 * rho * vel * vel)
    w = np.polynomial.legendre.leggauss(4)[1]
    fr, sr = MO.features(u.reshape(mesh.E, -1), detJ.reshape(mesh.E, 16), w, _hmin(mesh), inv.cpu().numpy(),
                         uref, 2, 4, m.nf)
    np.testing.assert_allclose(feat.cpu().numpy().reshape(mesh.E, m.nf), fr, rtol=2e-6, atol=2e-6)
    np.testing.assert_allclose(sig.cpu().numpy(), sr, rtol=2e-6)
    m.close()
    ctx.close()


@pytest.mark.parametrize("kind", KINDS)
def test_learned_preconditioner_trains_on_solver_blocks_and_converges(dev, kind):
    from dpgsolver_amd import _lib

    mesh, prim = _channel(16, 8)
    ctx = _ctx(mesh, prim)
    nb, E = 64, mesh.E
    u0 = M.uniform_state(mesh, prim)
    cap = E * 120
    m = _mlp(nb, 16, kind=kind)
    blocks = torch.empty(cap * nb * nb, dtype=torch.float64, device=dev)
    feat = torch.empty(cap * m.nf, dtype=torch.float32, device=dev)
    sig = torch.empty(cap, dtype=torch.float32, device=dev)

    # 1. exact LU block-Jacobi (the fallback and the method the learned one is measured against), capturing its blocks
    u_bj = torch.from_numpy(u0).to(dev)
    st_bj = ctx.solve_implicit(u_bj.data_ptr(), cfl_initial=1.0, cfl_max=300.0, exit_tol=1e-10, max_newton=120,
                               precond=_lib.PC_BLOCK_JACOBI,
                               capture=(blocks.data_ptr(), feat.data_ptr(), sig.data_ptr(), cap))
    n = st_bj["captured"]
    assert st_bj["converged"] == 1 and n == E * st_bj["newton_steps"]

    # 2. the untrained network (P = -sigma I): the method without the learned part
    ctx.attach_mlp(m)
    u_un = torch.from_numpy(u0).to(dev)
    st_un = ctx.solve_implicit(u_un.data_ptr(), cfl_initial=1.0, cfl_max=300.0, exit_tol=1e-10, max_newton=200,
                               precond=_lib.PC_LEARNED, raise_on_error=False)

    # 3. train on the captured blocks
    l0, _ = m.loss_grad(feat.data_ptr(), sig.data_ptr(), blocks.data_ptr(), min(n, 4096), want_grad=False)
    l1 = m.train(feat.data_ptr(), sig.data_ptr(), blocks.data_ptr(), n, epochs=30, batch=256, lr=1e-3)
    print(f"captured {n} blocks; loss untrained {l0:.4f} -> trained {l1:.4f}")
    assert l1 < 0.8 * l0

    # 4. the trained network inside GMRES; the diagonal + rank-8 structure is too weak to finish the solve
    #    alone, so it runs with the exact block-Jacobi fallback; the Kronecker structure finishes on its own
    u_ln = torch.from_numpy(u0).to(dev)
    fb = 100 if kind == 0 else 0
    st_ln = ctx.solve_implicit(u_ln.data_ptr(), cfl_initial=1.0, cfl_max=300.0, exit_tol=1e-10, max_newton=200,
                               precond=_lib.PC_LEARNED, learned_fallback_iters=fb)
    for name, st in (("block-jacobi", st_bj), ("untrained", st_un), ("learned", st_ln)):
        print(f"{name:13s} newton {st['newton_steps']:4d} gmres {st['gmres_iters']:6d} "
              f"max_rhs {st['max_rhs']:.2e} {st['seconds']:.3f} s fallbacks {st['fallback_steps']}")
    assert st_ln["converged"] == 1 and st_ln["max_rhs"] < 1e-10
    assert st_ln["gmres_iters"] < st_un["gmres_iters"]
    if kind == 1:
        assert st_ln["fallback_steps"] == 0
    a, b = u_ln.cpu().numpy(), u_bj.cpu().numpy()
    assert np.abs(a - b).max() < 1e-7 * np.abs(b).max()     # the same discrete steady state
    m.close()
    ctx.close()
