"""CPU oracle of the learned block preconditioner (float64 numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ (and nothing in the product path).  It restates, in plain
float64 linear algebra, the network, the loss and the analytic gradient that libdgx's HIP kernels
compute, so that the GPU forward pass, loss and gradient can be checked element by element.

The reference has no learned component.  What the network approximates is the element diagonal block
of the reference's pseudo-time Newton matrix, A_e = dR_e/du_e - M_e/dt_e: the blocks that
solvers/solve_implicit.c:232-256 (implicit_step) assembles through compute_rlhs_v / compute_rlhs_f and
hands to PETSc's block-Jacobi preconditioner (petsc_options_gmres_default.txt).  Its inverse is what an
exact block-Jacobi preconditioner applies.

Model (shared with dgx_mlp.hip):
  features x_e (nf)  ->  h1 = tanh(W1 x + b1)  ->  h2 = tanh(W2 h1 + b2)  ->  o = W3 h2 + b3
  kind 0 (low rank): o = [d (nb) | U (nb x R, row-major) | V (nb x R, row-major)],
                     P_e = sigma_e (diag(d) + U V^T);
  kind 1 (Kronecker): o = [S_0..S_{R-1} (nv x nv) | Q_0..Q_{R-1} (Np x Np)],
                     P_e = sigma_e sum_r S_r (x) Q_r  (rows variable-major, like the state);
  L = (1/B) sum_e ||P_e A_e - I||_F^2 / nb.
With ``emulate_bf16`` every GEMM operand is rounded to bfloat16 (round to nearest even) as the MFMA
kernels do, while accumulation stays float64; biases and tanh derivatives use the unrounded values, as
on the GPU.

Pins (tests/test_mlp_oracle.py):
  * the gradient against central finite differences of the float64 loss;
  * L = 1 exactly when the network outputs zero (P = 0, P A - I = -I);
  * L = 0 when A is the numpy inverse of sigma (diag(d) + U V^T), or of sigma sum_r np.kron(S_r, Q_r)
    (fixes the output layouts and sigma);
  * bf16 rounding against torch's float32 -> bfloat16 conversion.
"""
from __future__ import annotations

import numpy as np


def bf16_round(x):
    """Round float32 values to the nearest bfloat16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    finite = (u & 0x7F800000) != 0x7F800000
    r = np.where(finite, (u + 0x7FFF + ((u >> 16) & 1)) >> 16, u >> 16)
    out = ((r & 0xFFFF) << 16).astype(np.uint32).view(np.float32)
    return out.astype(np.float64).reshape(np.shape(x))


def layout(nf: int, nhid: int, nout: int):
    """Offsets and shapes of W1, b1, W2, b2, W3, b3 in the flat parameter vector."""
    shapes = [(nhid, nf), (nhid,), (nhid, nhid), (nhid,), (nout, nhid), (nout,)]
    offs, acc = [], 0
    for s in shapes:
        offs.append(acc)
        acc += int(np.prod(s))
    return offs, shapes, acc


def unpack(params, nf, nhid, nout):
    offs, shapes, n = layout(nf, nhid, nout)
    assert params.size == n
    return [np.asarray(params[o:o + int(np.prod(s))], dtype=np.float64).reshape(s) for o, s in zip(offs, shapes)]


def _r(x, emulate):
    return bf16_round(x) if emulate else np.asarray(x, dtype=np.float64)


def forward(params, X, nf, nhid, nout, emulate_bf16=False):
    """Return (h1, h2, o) for a batch X (B x nf)."""
    W1, b1, W2, b2, W3, b3 = unpack(_r(params, emulate_bf16), nf, nhid, nout)
    e = emulate_bf16
    h1 = np.tanh(_r(X, e) @ W1.T + b1)
    h2 = np.tanh(_r(h1, e) @ W2.T + b2)
    o = _r(h2, e) @ W3.T + b3
    return h1, h2, o


def split_output(o, nb, R):
    d = o[:, :nb]
    U = o[:, nb:nb + nb * R].reshape(-1, nb, R)
    V = o[:, nb + nb * R:nb + 2 * nb * R].reshape(-1, nb, R)
    return d, U, V


def n_out(nb, R, kind=0, Np=None):
    if kind == 0:
        return nb + 2 * nb * R
    nv = nb // Np
    return R * (nv * nv + Np * Np)


def split_kron(o, nb, R, Np):
    nv = nb // Np
    S = o[:, :R * nv * nv].reshape(-1, R, nv, nv)
    Q = o[:, R * nv * nv:].reshape(-1, R, Np, Np)
    return S, Q


def precond_blocks(o, sigma, nb, R, kind=0, Np=None):
    """P_e of the chosen structure, shape (B, nb, nb)."""
    if kind == 0:
        d, U, V = split_output(o, nb, R)
        P = np.einsum("bir,bjr->bij", U, V)
        P[:, np.arange(nb), np.arange(nb)] += d
    else:
        S, Q = split_kron(o, nb, R, Np)
        P = np.zeros((len(o), nb, nb))
        for e in range(len(o)):
            for r in range(R):
                P[e] += np.kron(S[e, r], Q[e, r])
    return P * np.asarray(sigma, dtype=np.float64)[:, None, None]


def block_loss(o, sigma, A, nb, R, kind=0, Np=None):
    """Per-sample L_e = ||P_e A_e - I||_F^2 / nb and dL_e/do (B x nout)."""
    sg = np.asarray(sigma, dtype=np.float64)
    P = precond_blocks(o, sg, nb, R, kind, Np)
    Rm = P @ A - np.eye(nb)[None]
    L = (Rm ** 2).sum(axis=(1, 2)) / nb
    G = 2.0 / nb * Rm @ np.transpose(A, (0, 2, 1)) * sg[:, None, None]    # sigma dL/dP
    if kind == 0:
        d, U, V = split_output(o, nb, R)
        go = np.concatenate([np.diagonal(G, axis1=1, axis2=2),
                             (G @ V).reshape(len(o), -1),
                             (np.transpose(G, (0, 2, 1)) @ U).reshape(len(o), -1)], axis=1)
    else:
        nv = nb // Np
        S, Q = split_kron(o, nb, R, Np)
        G4 = G.reshape(len(o), nv, Np, nv, Np)                # [e, a, i, b, j]
        dS = np.einsum("eaibj,erij->erab", G4, Q)
        dQ = np.einsum("eaibj,erab->erij", G4, S)
        go = np.concatenate([dS.reshape(len(o), -1), dQ.reshape(len(o), -1)], axis=1)
    return L, go


def loss_and_grad(params, X, sigma, A, nb, R, nhid, emulate_bf16=False, kind=0, Np=None):
    """Mean loss over the batch and its gradient with respect to the flat parameter vector."""
    B, nf = X.shape
    nout = n_out(nb, R, kind, Np)
    e = emulate_bf16
    Wb = unpack(_r(params, e), nf, nhid, nout)
    W2, W3 = Wb[2], Wb[4]
    h1, h2, o = forward(params, X, nf, nhid, nout, e)
    L, go = block_loss(o, sigma, A, nb, R, kind, Np)
    go = go / B
    gW3 = _r(go, e).T @ _r(h2, e)
    gb3 = go.sum(0)
    gh2 = (_r(go, e) @ W3) * (1.0 - h2 ** 2)
    gW2 = _r(gh2, e).T @ _r(h1, e)
    gb2 = gh2.sum(0)
    gh1 = (_r(gh2, e) @ W2) * (1.0 - h1 ** 2)
    gW1 = _r(gh1, e).T @ _r(X, e)
    gb1 = gh1.sum(0)
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3)])
    return L.mean(), g


def features(u, detJ, w_gl, hmin, inv_dt, u_ref, d, N, nf):
    """Element features and the analytic scale sigma_e for 2-D Euler (nv = 4).

    u: (E, 4*Np) conserved coefficients; detJ: (E, Np) at the tensor Gauss-Legendre nodes;
    w_gl: 1-D Gauss-Legendre weights; hmin, inv_dt: (E,); u_ref: (rho, momentum, energy) scales.
    Returns feat (E, nf) = [u / u_ref | detJ / mean(detJ) | log10(h inv_dt / lambda) | 0] and
    sigma (E,) = 1 / (mean(w detJ) inv_dt + h lambda / 4), lambda = max_nodes(|v| + c).
    """
    E = u.shape[0]
    Np = N ** d
    W = np.kron(w_gl, w_gl) if d == 2 else np.kron(np.kron(w_gl, w_gl), w_gl)
    uu = u.reshape(E, d + 2, Np)
    rho = uu[:, 0]
    V2 = sum((uu[:, 1 + k] / rho) ** 2 for k in range(d))
    p = 0.4 * (uu[:, d + 1] - 0.5 * rho * V2)
    lam = (np.sqrt(V2) + np.sqrt(np.maximum(1.4 * p / rho, 0.0))).max(1)
    mwj = (W[None] * detJ).sum(1) / Np
    sigma = 1.0 / (mwj * inv_dt + 0.25 * hmin * lam)
    ref = np.array([u_ref[0]] + [u_ref[1]] * d + [u_ref[2]])
    feat = np.zeros((E, nf))
    nb = (d + 2) * Np
    feat[:, :nb] = (uu / ref[None, :, None]).reshape(E, nb)
    feat[:, nb:nb + Np] = detJ / detJ.mean(1, keepdims=True)
    feat[:, nb + Np] = np.log10(hmin * inv_dt / lam + 1e-30)
    return feat, sigma
