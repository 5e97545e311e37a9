"""Pins of the learned-preconditioner oracle (CPU only).

The oracle's analytic gradient is checked against central finite differences of its own float64 loss,
the loss against two closed forms (zero output; the exact inverse), and its bf16 rounding against
torch's float32 -> bfloat16 conversion.  A planted mistake (a transposed U/V, a dropped sigma, a wrong
tanh derivative) fails one of these.
"""
import numpy as np
import pytest

import mlp_oracle as MO


def _rand_problem(rng, B=5, nf=7, nhid=6, nb=5, R=2, kind=0, Np=None):
    nout = MO.n_out(nb, R, kind, Np)
    _, _, n = MO.layout(nf, nhid, nout)
    params = rng.standard_normal(n) * 0.4
    X = rng.standard_normal((B, nf))
    A = rng.standard_normal((B, nb, nb)) + 3 * np.eye(nb)
    sigma = rng.uniform(0.5, 2.0, B)
    return params, X, sigma, A, nb, R, nhid


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(11)
    params, X, sigma, A, nb, R, nhid = _rand_problem(rng)
    L0, g = MO.loss_and_grad(params, X, sigma, A, nb, R, nhid)
    h = 1e-6
    fd = np.zeros_like(g)
    for i in range(params.size):
        e = np.zeros_like(params)
        e[i] = h
        fd[i] = (MO.loss_and_grad(params + e, X, sigma, A, nb, R, nhid)[0]
                 - MO.loss_and_grad(params - e, X, sigma, A, nb, R, nhid)[0]) / (2 * h)
    assert np.abs(g - fd).max() <= 1e-7 * max(1.0, np.abs(fd).max())


def test_kron_gradient_matches_finite_differences():
    rng = np.random.default_rng(17)
    params, X, sigma, A, nb, R, nhid = _rand_problem(rng, nb=6, R=2, kind=1, Np=3)
    kw = dict(kind=1, Np=3)
    _, g = MO.loss_and_grad(params, X, sigma, A, nb, R, nhid, **kw)
    h = 1e-6
    fd = np.zeros_like(g)
    for i in range(params.size):
        e = np.zeros_like(params)
        e[i] = h
        fd[i] = (MO.loss_and_grad(params + e, X, sigma, A, nb, R, nhid, **kw)[0]
                 - MO.loss_and_grad(params - e, X, sigma, A, nb, R, nhid, **kw)[0]) / (2 * h)
    assert np.abs(g - fd).max() <= 1e-7 * max(1.0, np.abs(fd).max())


def test_kron_exact_inverse_gives_zero_loss():
    rng = np.random.default_rng(18)
    nv, Np, R = 2, 3, 2
    nb = nv * Np
    S = rng.standard_normal((R, nv, nv)) + 2 * np.eye(nv)
    Q = rng.standard_normal((R, Np, Np)) * 0.2 + np.eye(Np)
    o = np.concatenate([S.ravel(), Q.ravel()])[None]
    P = 1.3 * sum(np.kron(S[r], Q[r]) for r in range(R))
    A = np.linalg.inv(P)[None]
    L, go = MO.block_loss(o, np.array([1.3]), A, nb, R, kind=1, Np=Np)
    assert L[0] < 1e-22 and np.abs(go).max() < 1e-10
    # the swapped factor order (Q (x) S) is a different matrix
    Pt = 1.3 * sum(np.kron(Q[r], S[r]) for r in range(R))
    assert ((Pt @ A[0] - np.eye(nb)) ** 2).sum() / nb > 1e-4


def test_gradient_plants_fail():
    """The finite-difference pin catches a wrong tanh derivative planted in the oracle."""
    rng = np.random.default_rng(12)
    params, X, sigma, A, nb, R, nhid = _rand_problem(rng)
    _, g = MO.loss_and_grad(params, X, sigma, A, nb, R, nhid)
    # recompute with the derivative of tanh taken as 1 (a plausible mistake)
    B, nf = X.shape
    nout = nb + 2 * nb * R
    W1, b1, W2, b2, W3, b3 = MO.unpack(params, nf, nhid, nout)
    h1, h2, o = MO.forward(params, X, nf, nhid, nout)
    _, go = MO.block_loss(o, sigma, A, nb, R)
    go /= B
    gh2 = go @ W3           # missing (1 - h2^2)
    gW2_bad = gh2.T @ h1
    offs, _, _ = MO.layout(nf, nhid, nout)
    gW2 = g[offs[2]:offs[3]].reshape(nhid, nhid)
    assert np.abs(gW2 - gW2_bad).max() > 1e-3


def test_zero_output_gives_unit_loss():
    rng = np.random.default_rng(13)
    params, X, sigma, A, nb, R, nhid = _rand_problem(rng)
    nf = X.shape[1]
    nout = nb + 2 * nb * R
    offs, _, _ = MO.layout(nf, nhid, nout)
    params[offs[4]:] = 0.0       # W3 = 0, b3 = 0
    L, _ = MO.loss_and_grad(params, X, sigma, A, nb, R, nhid)
    assert L == 1.0


def test_exact_inverse_gives_zero_loss():
    rng = np.random.default_rng(14)
    nb, R = 6, 2
    d = rng.uniform(1, 2, nb)
    U = rng.standard_normal((nb, R)) * 0.3
    V = rng.standard_normal((nb, R)) * 0.3
    o = np.concatenate([d, U.ravel(), V.ravel()])[None]
    sigma = np.array([0.7])
    P = 0.7 * (np.diag(d) + U @ V.T)
    A = np.linalg.inv(P)[None]
    L, go = MO.block_loss(o, sigma, A, nb, R)
    assert L[0] < 1e-24 and np.abs(go).max() < 1e-12
    # a transposed V (V U^T instead of U V^T) is not the inverse
    Pt = 0.7 * (np.diag(d) + V @ U.T)
    assert ((Pt @ A[0] - np.eye(nb)) ** 2).sum() / nb > 1e-4


def test_bf16_round_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(15)
    x = np.concatenate([rng.standard_normal(4000).astype(np.float32) * 10.0 ** rng.integers(-6, 6, 4000),
                        np.float32([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -1 - 2 ** -8, 0.0, -0.0, 3.0e38])])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(MO.bf16_round(x), ref)
    assert MO.bf16_round(np.float32([1 + 2 ** -8]))[0] == 1.0
    assert MO.bf16_round(np.float32([1 + 3 * 2 ** -8]))[0] == 1 + 2 ** -6


def test_emulated_bf16_close_to_float64():
    rng = np.random.default_rng(16)
    params, X, sigma, A, nb, R, nhid = _rand_problem(rng)
    L64, g64 = MO.loss_and_grad(params, X, sigma, A, nb, R, nhid)
    L16, g16 = MO.loss_and_grad(params, X, sigma, A, nb, R, nhid, emulate_bf16=True)
    assert abs(L16 - L64) < 0.05 * L64
    assert np.linalg.norm(g16 - g64) < 0.05 * np.linalg.norm(g64)


def test_features_uniform_state_closed_form():
    E, d, N = 3, 2, 4
    Np = N * N
    rho, ux, uy, p = 1.2, 0.3, -0.1, 0.9
    E_tot = p / 0.4 + 0.5 * rho * (ux * ux + uy * uy)
    u = np.tile(np.repeat([rho, rho * ux, rho * uy, E_tot], Np), (E, 1))
    w = np.polynomial.legendre.leggauss(N)[1]
    detJ = np.full((E, Np), 0.25)
    hmin = np.array([0.1, 0.2, 0.3])
    inv_dt = np.array([1.0, 10.0, 100.0])
    feat, sigma = MO.features(u, detJ, w, hmin, inv_dt, (1.0, 1.0, 2.0), d, N, 96)
    lam = np.hypot(ux, uy) + np.sqrt(1.4 * p / rho)
    # sum of the 2-D weights is 4: mean(w detJ) = 4 * 0.25 / 16
    np.testing.assert_allclose(sigma, 1 / (0.0625 * inv_dt + 0.25 * hmin * lam), rtol=1e-14)
    np.testing.assert_allclose(feat[:, 64:80], 1.0)
    np.testing.assert_allclose(feat[:, 80], np.log10(hmin * inv_dt / lam), rtol=1e-12)
    np.testing.assert_allclose(feat[0, 48:64], E_tot / 2.0)
    assert np.all(feat[:, 81:] == 0)
