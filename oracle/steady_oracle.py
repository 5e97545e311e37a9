"""steady_oracle.py -- TEST INFRASTRUCTURE ONLY (the CPU baseline of the steady solve; never part of the product).

The pseudo-time Newton-GMRES of the reference's implicit solver (solve_implicit.c:95-110, 232-281, 393-470,
780-873) on the CPU, on the oracle's FP64 residual, in the form the GPU driver dgx_solve_implicit runs it:

* per Newton step: the residual R(u) and max |R| (exit when max |R| < exit_tol); the ramped CFL number
  cfl_initial * max(1, max|R_0| / max|R|), capped at cfl_max (compute_CFL_ramping, solve_dg.c:223-248); the
  per-element 1/dt of the CFL-constrained step (or_cfl_inv_dt); the operator A = dR/du - M/dt;
* the preconditioner: exact element diagonal blocks of A (complex step, or_diag_blocks_cs_threads) inverted
  (Gauss-Jordan with partial pivoting, or_block_inverse), applied block by block (block-Jacobi, the fallback and
  judge of the learned preconditioner);
* right-preconditioned restarted GMRES(m) with classical Gram-Schmidt applied twice and Givens rotations, x0 = 0,
  stopped at |r| <= rtol |b| (the residual estimate checked every step) or max_iter iterations; A v by complex
  step on the threaded complex residual (or_compute_rhs_c_threads) minus the mass term;
* the relaxed update u += relax_e x_e (or_relaxed_update) and the positivity limiter (or_enforce_positivity).

Every residual, block and vector product runs on nthreads OpenMP threads (numpy's own BLAS threads should be
limited to one, so that the two thread pools do not compete); the wall time includes everything from the first
residual to the converged one.
"""
import time

import numpy as np

import pyoracle


def _gmres(A, P, b, m, rtol, max_iter, nthreads):
    n = b.size
    x = np.zeros(n)
    bnorm = np.linalg.norm(b)
    if bnorm == 0.0:
        return x, 0, True
    tol = rtol * bnorm
    total = 0
    V = np.empty((m + 1, n))
    while total < max_iter:
        r = b - A(x) if total else b.copy()
        beta = np.linalg.norm(r)
        if beta <= tol:
            return x, total, True
        V[0] = r / beta
        H = np.zeros((m + 1, m))
        cs, sn = np.zeros(m), np.zeros(m)
        g = np.zeros(m + 1)
        g[0] = beta
        conv = False
        j = 0
        while j < m and total < max_iter:
            w = A(P(V[j]))
            h = pyoracle.dots(V, j + 1, w, nthreads)
            pyoracle.gemv_sub(V, j + 1, h, w, nthreads)
            h2 = pyoracle.dots(V, j + 1, w, nthreads)
            pyoracle.gemv_sub(V, j + 1, h2, w, nthreads)
            H[: j + 1, j] = h + h2
            H[j + 1, j] = np.linalg.norm(w)
            V[j + 1] = w / H[j + 1, j]
            for i in range(j):
                t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
                H[i, j] = t
            den = np.hypot(H[j, j], H[j + 1, j])
            cs[j], sn[j] = H[j, j] / den, H[j + 1, j] / den
            H[j, j] = den
            H[j + 1, j] = 0.0
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j] * g[j]
            j += 1
            total += 1
            if abs(g[j]) <= tol:
                conv = True
                break
        y = np.zeros(j)
        for i in range(j - 1, -1, -1):
            y[i] = (g[i] - H[i, i + 1: j] @ y[i + 1: j]) / H[i, i]
        x = x + P(y @ V[:j])
        if conv:
            return x, total, True
    return x, total, False


def solve_implicit(orc, u, nthreads, cfl_initial=1.0, cfl_max=300.0, exit_tol=1e-10, max_newton=100, restart=60,
                   rtol=1e-4, max_iter=1000, h=1e-30, progress=None):
    """Runs the solve in place on u (float64, contiguous); returns the statistics and the wall time.
    progress(step, max_rhs, gmres_iterations, seconds) is called after each Newton step."""
    mesh = orc.mesh
    E, Np, nv = mesh.E, orc.Np, orc.nv
    nb = Np * nv
    mass = np.stack([orc.mass(e) for e in range(E)])                  # (E, Np, Np)
    eye = np.eye(nv)
    stats = {"newton_steps": 0, "gmres_iters": 0, "gmres_per_step": [], "history": [], "converged": False}
    t0 = time.perf_counter()
    mx0 = None
    for it in range(max_newton):
        r = orc.rhs_threads(u, nthreads)
        mx = float(np.abs(r).max())
        stats["history"].append(mx)
        mx0 = mx if mx0 is None else mx0
        stats["max_rhs"] = mx
        if mx < exit_tol:
            stats["converged"] = True
            break
        cfl = min(cfl_initial * max(1.0, mx0 / mx), cfl_max) if cfl_max > 0 else cfl_initial * max(1.0, mx0 / mx)
        inv_dt = orc.cfl_inv_dt(u, cfl)
        blocks = orc.diag_blocks(u, h, nthreads)
        blocks -= np.einsum("vw,eij,e->eviwj", eye, mass, inv_dt).reshape(E, nb, nb)
        inv = pyoracle.block_inverse(blocks, nthreads)
        mdt = mass * inv_dt[:, None, None]

        def A(v):
            _, jim = orc.rhs_complex_threads(u, h * v, nthreads)
            return jim / h - np.matmul(mdt[:, None], v.reshape(E, nv, Np, 1)).reshape(-1)

        P = lambda v: pyoracle.block_apply(inv, v, nthreads)
        x, its, conv = _gmres(A, P, -r, restart, rtol, max_iter, nthreads)
        stats["gmres_per_step"].append(its)
        stats["gmres_iters"] += its
        if orc.relaxed_update(u, x):
            raise RuntimeError("relaxation below EPS")
        limited, status = orc.positivity(u)
        if status:
            raise RuntimeError("positivity limiter failed")
        u[:] = limited
        stats["newton_steps"] = it + 1
        if progress:
            progress(it, mx, its, time.perf_counter() - t0)
    stats["seconds"] = time.perf_counter() - t0
    return stats
