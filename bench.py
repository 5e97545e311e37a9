"""Residual-evaluation throughput of the MI355X DG path (GDOF/s), one JSON line on stdout.

Workload: 3-D Euler, DG p=4 on a 64^3 periodic HEX mesh of [0,1]^3 (262,144 elements, 125 nodes x
5 variables each, 163.84 M DOF), curved by x += 0.05 sin(2 pi y) sin(2 pi z) (cyclic),
Lax-Friedrichs flux, seeded smooth state rho in [0.9,1.1], |u| <= 0.485, p in [0.9,1.1].
A "step" is one residual evaluation R(u) (volume + face terms, compute_rlhs_common_dg) with the
state already resident in HBM, launched as one replay of a HIP graph that holds the residual kernel and
timed with HIP events on the launch stream.  `value` is the mixed residual (FP64 state in, FP32 contractions,
FP64 pointwise fluxes, FP32 residual out), the FP32 inner residual whose error against the FP64 residual stays
below 1e-6 (rel_err_vs_fp64 beside it); the FP64 residual (the outer residual, 1e-12 against the oracle in the
tests) and the plain FP32 residual (FP32 state, error above 1e-6) are timed beside it.

CPU baseline: the FP64 oracle (a C restatement of the reference's dense per-element operator
application) on one host core, timed on a bounded sample of the same workload (an 8^3 periodic
mesh of the same curved HEX p=4 elements), and the same on all host cores (OpenMP, at most 16).  For the
steady channel solve, the oracle's Newton-GMRES with LU block-Jacobi on one core, estimated from a bounded
sample (cpu_steady_estimate).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def _state_on_device(torch, xyz_dev, seed):
    """config3_state (mesh.py) evaluated with torch on the device; returns [E][5][Np] float64."""
    from dpgsolver_amd import mesh as M

    rng = M.SplitMix64(seed + 3)
    phi = rng.uniform(15, 0.0, 1.0).reshape(5, 3)
    s = []
    sx = [torch.sin(2.0 * math.pi * xyz_dev[:, c, :]) for c in range(3)]
    cx = [torch.cos(2.0 * math.pi * xyz_dev[:, c, :]) for c in range(3)]
    for k in range(5):
        v = None
        for c in range(3):
            t = sx[c] * math.cos(2.0 * math.pi * phi[k, c]) + cx[c] * math.sin(2.0 * math.pi * phi[k, c])
            v = t if v is None else v * t
        s.append(v)
    rho = 1.0 + 0.1 * s[0]
    vel = [0.28 * s[1 + c] for c in range(3)]
    p = 1.0 + 0.1 * s[4]
    E, Np = xyz_dev.shape[0], xyz_dev.shape[2]
    out = torch.empty((E, 5, Np), dtype=torch.float64, device=xyz_dev.device)
    out[:, 0] = rho
    ke = torch.zeros_like(rho)
    for c in range(3):
        out[:, 1 + c] = rho * vel[c]
        ke += vel[c] * vel[c]
    out[:, 4] = p / M.GM1 + 0.5 * rho * ke
    return out.reshape(-1)


def _traffic_from_profiles(kernel_tag, summary="rhs_hex_p4_64cube_traffic.md"):
    """HBM bytes per launch (calibrated rocprofv3 FETCH_SIZE + WRITE_SIZE) from a committed summary, if present:
    lines `traffic <tag> <bytes>` in profiles/<summary>."""
    path = os.path.join(ROOT, "profiles", summary)
    try:
        for line in open(path):
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic" and parts[1] == kernel_tag:
                return int(parts[2])
    except (OSError, ValueError):
        return None
    return None


def cpu_baseline(target_seconds=12.0):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from dpgsolver_amd import mesh as M

    mesh = M.config3_mesh(8, 4)
    u = M.config3_state(mesh)
    orc = pyoracle.Oracle(mesh, M.PDE_EULER, M.FLUX_LF)
    orc.rhs(u)
    n, t0 = 0, time.perf_counter()
    while True:
        orc.rhs(u)
        n += 1
        el = time.perf_counter() - t0
        if el >= target_seconds or n >= 200:
            break
    per = el / n
    dof = mesh.E * 125 * 5
    out = {"value": dof / per / 1e9, "unit": "GDOF/s", "cores": 1, "kind": "port",
           "sample": f"oracle dense-operator FP64 residual, 8^3 curved HEX p=4 periodic mesh ({mesh.E} elements), "
                     f"{n} evaluations in {el:.1f} s on 1 host core"}
    # all host cores (OpenMP over elements, faces lifted by colour), capped at the 16 threads of a 1-GPU share
    host_cores = len(os.sched_getaffinity(0))
    cores = max(1, min(host_cores, 16))
    orc.rhs_threads(u, cores)
    n2, t0 = 0, time.perf_counter()
    while True:
        orc.rhs_threads(u, cores)
        n2 += 1
        el2 = time.perf_counter() - t0
        if el2 >= target_seconds / 3 or n2 >= 400:
            break
    out["all_cores"] = {"value": dof / (el2 / n2) / 1e9, "unit": "GDOF/s", "cores": cores, "kind": "port",
                        "host_cores_available": host_cores,
                        "sample": f"same residual on {cores} OpenMP threads ({host_cores} cores in this process's "
                                  f"affinity mask; a one-GPU share is 16), {n2} evaluations in {el2:.1f} s"}
    return out


def cpu_steady_estimate(newton_steps, gmres_iters, restart=60, n_blocks=1024, n_iters=30):
    """CPU time-to-1e-10 of the same Newton-GMRES with exact LU block-Jacobi, estimated from a bounded sample on
    one host core: the oracle's complex-step diagonal blocks of n_blocks elements (A_e = dR_e/du_e - M_e/dt)
    and their inverses, scaled to all elements, per Newton step; and n_iters right-preconditioned GMRES
    iterations (block-inverse apply, complex-step J.v minus the mass term, two classical Gram-Schmidt passes
    over a basis that grows to n_iters vectors), per GMRES iteration.  The estimate is
    newton_steps x (residual + set-up) + gmres_iters x (one iteration at the average basis size of GMRES(60)),
    with the Newton and GMRES counts of the GPU block-Jacobi leg."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle
    from threadpoolctl import threadpool_limits
    from dpgsolver_amd import mesh as M

    mesh = M.channel_mesh(64, 64, 3)
    prim = M.freestream_prim(2)
    orc = pyoracle.Oracle(mesh, M.PDE_EULER, M.FLUX_ROE, bc_prim=prim)
    u = M.uniform_state(mesh, prim) * (1.0 + 1e-3 * np.random.default_rng(5).uniform(-1, 1, mesh.E * 64))
    E, nb, n = mesh.E, 64, mesh.E * 64
    dt = 0.1
    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        orc.rhs(u)
        t_res = time.perf_counter() - t0
        idx = np.random.default_rng(0).choice(E, n_blocks, replace=False)
        t0 = time.perf_counter()
        blk = np.empty((n_blocks, nb, nb))
        for k, e in enumerate(idx):
            blk[k] = orc.jacobian_block(u, int(e), int(e)) - np.kron(np.eye(4), orc.mass(int(e))) / dt
        inv = np.linalg.inv(blk)
        t_setup = (time.perf_counter() - t0) * E / n_blocks
        inv_all = np.tile(inv, (E // n_blocks + 1, 1, 1))[:E]
        mass = np.stack([orc.mass(e) for e in range(E)])
        V = np.zeros((n_iters + 1, n))
        r0 = np.random.default_rng(1).standard_normal(n)
        V[0] = r0 / np.linalg.norm(r0)
        t0 = time.perf_counter()
        for j in range(n_iters):
            z = np.einsum("eij,ej->ei", inv_all, V[j].reshape(E, nb)).reshape(-1)
            jre, jim = orc.rhs_complex(u, 1e-30 * z)
            w = jim / 1e-30 - np.einsum("eab,evb->eva", mass, z.reshape(E, 4, 16)).reshape(-1) / dt
            for _ in range(2):
                h = V[: j + 1] @ w
                w -= h @ V[: j + 1]
            V[j + 1] = w / np.linalg.norm(w)
        t_it = (time.perf_counter() - t0) / n_iters
    # CGS2 grows linearly with the basis: the sampled iterations average (n_iters+1)/2 basis vectors, GMRES(restart)
    # averages (restart+1)/2; add the difference at the measured cost of CGS2 over n_iters vectors
    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        w = V[n_iters].copy()
        for _ in range(2):
            h = V[:n_iters] @ w
            w -= h @ V[:n_iters]
        t_cgs = time.perf_counter() - t0
    t_it_full = t_it + t_cgs * ((restart + 1) / 2 - (n_iters + 1) / 2) / n_iters
    total = newton_steps * (t_res + t_setup) + gmres_iters * t_it_full
    return {"seconds_estimate": round(total, 1), "cores": 1, "kind": "port",
            "per_newton_setup_s": round(t_res + t_setup, 3), "per_gmres_iter_s": round(t_it_full, 5),
            "newton_steps": newton_steps, "gmres_iters": gmres_iters,
            "sample": f"oracle complex-step diagonal blocks + inverses of {n_blocks} of {E} elements (scaled), "
                      f"{n_iters} GMRES iterations (complex-step J.v, block-inverse apply, CGS2) on 1 core; "
                      "counts from the GPU block-Jacobi solve"}


def cpu_steady_measured(gpu_compare=True):
    """The same steady solve measured on the host's cores: the oracle's restatement of the pseudo-time
    Newton-GMRES(60) with exact LU block-Jacobi (oracle/steady_oracle.py: threaded residual, complex-step J.v and
    diagonal blocks, Gauss-Jordan block inverses, CGS2), with the parameters of the GPU block-Jacobi leg, from the
    free stream to max|R| < 1e-10, on min(16, available) cores.  Then (gpu_compare) the GPU block-Jacobi solve of
    the same problem, and the difference of the two converged states."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle
    import steady_oracle
    from threadpoolctl import threadpool_limits
    from dpgsolver_amd import mesh as M

    avail = len(os.sched_getaffinity(0))
    nt = min(16, avail)
    mesh = M.channel_mesh(64, 64, 3)
    prim = M.freestream_prim(2)
    orc = pyoracle.Oracle(mesh, M.PDE_EULER, M.FLUX_ROE, bc_prim=prim)
    u = np.ascontiguousarray(M.uniform_state(mesh, prim))
    print(f"cpu steady solve on {nt} threads ({avail} cores available)", file=sys.stderr, flush=True)
    with threadpool_limits(limits=1):
        st = steady_oracle.solve_implicit(
            orc, u, nt, cfl_initial=1.0, cfl_max=300.0, exit_tol=1e-10, max_newton=100, restart=60, rtol=1e-4,
            max_iter=1000, progress=lambda it, mx, its, t: print(f"  cpu newton {it}: max|R| {mx:.3e}, gmres {its}, "
                                                                 f"{t:.1f} s", file=sys.stderr, flush=True))
    print(f"cpu block-Jacobi: {st['seconds']:.1f} s, newton {st['newton_steps']}, gmres {st['gmres_iters']}",
          file=sys.stderr, flush=True)
    out = {"workload": "2-D Euler DG p=3, 64x64 curved-bump channel, Mach 0.5, Roe; Newton-GMRES(60) to 1e-10",
           "cpu": {"seconds": round(st["seconds"], 2), "newton_steps": st["newton_steps"],
                   "gmres_iters": st["gmres_iters"], "converged": st["converged"], "max_rhs": st["max_rhs"],
                   "gmres_per_step": st["gmres_per_step"], "cores": nt, "host_cores_available": avail,
                   "kind": "port",
                   "what": "oracle restatement of solve_implicit with exact LU block-Jacobi, every residual, "
                           "complex-step product, block and Gram-Schmidt pass on OpenMP threads"}}
    if gpu_compare:
        import torch
        from dpgsolver_amd import _lib
        dev = torch.device("cuda:0")
        ctx = _lib.Context(2, 3, M.PDE_EULER, M.FLUX_ROE, bc_prim=prim)
        ctx.set_stream(torch.cuda.current_stream(dev).cuda_stream)
        ctx.upload_mesh(mesh.xyz, mesh.faces)
        kw = dict(precond=_lib.PC_BLOCK_JACOBI, cfl_initial=1.0, cfl_max=300.0, exit_tol=1e-10, max_newton=100,
                  gmres_restart=60, gmres_rtol=1e-4, gmres_max_iter=1000, raise_on_error=False)
        ug = torch.from_numpy(M.uniform_state(mesh, prim)).to(dev)
        ctx.solve_implicit(ug.data_ptr(), **{**kw, "max_newton": 2})     # warm-up
        ug = torch.from_numpy(M.uniform_state(mesh, prim)).to(dev)
        sg = ctx.solve_implicit(ug.data_ptr(), **kw)
        ctx.close()
        ugh = ug.cpu().numpy()
        out["gpu"] = {"seconds": round(sg["seconds"], 4), "newton_steps": sg["newton_steps"],
                      "gmres_iters": sg["gmres_iters"], "converged": bool(sg["converged"]),
                      "gmres_per_step": sg["gmres_per_step"]}
        out["steady_states_rel_diff"] = float(np.abs(ugh - u).max() / np.abs(u).max())
        out["cpu_vs_gpu_wall"] = round(st["seconds"] / sg["seconds"], 1)
    return out


def time_to_steady(torch, dev, epochs=60, cpu=True):
    """Time to max|R| < 1e-10 (FP64) on the steady channel workload: 2-D Euler, DG p=3, 64x64 QUAD channel
    [-1.5,1.5]x[0,0.8] with the bump y = 0.0625 exp(-25 x^2), inflow Mach 0.5, Roe flux, pseudo-time
    Newton-GMRES(60) (rtol 1e-4) from the free stream with cfl_initial 1 and the ramped CFL capped at 300.
    Legs: exact block-Jacobi (FP64 inverse of every diagonal block), with the GMRES operator J.v - M v/dt in FP64
    and in FP32 dual numbers (the inner defect-correction product; the FP64 residual decides convergence); the
    learned preconditioner (Kronecker
    rank 8, trained here on the diagonal blocks captured from the block-Jacobi run) with the exact
    block-Jacobi as fallback when GMRES has not converged after 50 iterations; multicolour block symmetric
    Gauss-Seidel over the diagonal and face blocks, with the blocks stored in FP64 and in FP32."""
    from dpgsolver_amd import _lib, mesh as M

    mesh = M.channel_mesh(64, 64, 3)
    prim = M.freestream_prim(2)
    ctx = _lib.Context(2, 3, M.PDE_EULER, M.FLUX_ROE, bc_prim=prim)
    ctx.set_stream(torch.cuda.current_stream(dev).cuda_stream)
    ctx.upload_mesh(mesh.xyz, mesh.faces)
    u0 = M.uniform_state(mesh, prim)
    kw = dict(cfl_initial=1.0, cfl_max=300.0, exit_tol=1e-10, max_newton=100, gmres_restart=60, gmres_rtol=1e-4,
              gmres_max_iter=1000, raise_on_error=False)
    nb, E = 64, mesh.E
    mlp = _lib.MLP(nb, 16, hidden=256, rank=8, kind=_lib.MLP_KRON)
    cap = E * 40
    blocks = torch.empty(cap * nb * nb, dtype=torch.float64, device=dev)
    feat = torch.empty(cap * mlp.nf, dtype=torch.float32, device=dev)
    sig = torch.empty(cap, dtype=torch.float32, device=dev)

    def solve(pc, **extra):
        u = torch.from_numpy(u0).to(dev)
        st = ctx.solve_implicit(u.data_ptr(), precond=pc, **{**kw, **extra})
        return st, u

    def _training_mfma(n, epochs, seconds, batch=256, nf=96, nh=256, no=2176):
        # BF16 MFMA GEMMs per mini-batch: forward (3 layers), weight gradients (3) and input gradients (2)
        fwd = 2 * batch * (nf * nh + nh * nh + nh * no)
        step = 2 * fwd + 2 * batch * (no * nh + nh * nh)
        flop = step * epochs * ((n + batch - 1) // batch)
        ach = flop / seconds / 1e12
        return {"gemm_tflop": round(flop / 1e12, 3), "achieved_tflops": round(ach, 3), "peak_tflops": 2500.0,
                "frac": round(ach / 2500.0, 6),
                "note": "GEMM FLOPs over the whole training wall time (loss, bias-gradient and Adam kernels included)"}

    def leg(st):
        return {"seconds": round(st["seconds"], 4), "newton_steps": st["newton_steps"],
                "gmres_iters": st["gmres_iters"], "converged": bool(st["converged"]), "max_rhs": st["max_rhs"],
                "fallback_steps": st["fallback_steps"], "gmres_per_step": st["gmres_per_step"]}

    solve(_lib.PC_BLOCK_JACOBI, max_newton=2)      # warm-up (module load, first launches)
    st_cap, _ = solve(_lib.PC_BLOCK_JACOBI, capture=(blocks.data_ptr(), feat.data_ptr(), sig.data_ptr(), cap))
    n = st_cap["captured"]
    st_bj, u_bj = solve(_lib.PC_BLOCK_JACOBI)
    print(f"block-Jacobi: {st_bj['seconds']:.3f} s, newton {st_bj['newton_steps']}, gmres {st_bj['gmres_iters']}",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    l0, _ = mlp.loss_grad(feat.data_ptr(), sig.data_ptr(), blocks.data_ptr(), min(n, 4096), want_grad=False)
    loss = l0
    for ep in range(0, epochs, 5):
        lr = 1e-3 if ep < epochs // 3 else (3e-4 if ep < 2 * epochs // 3 else 1e-4)
        loss = mlp.train(feat.data_ptr(), sig.data_ptr(), blocks.data_ptr(), n, epochs=5, batch=256, lr=lr)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    print(f"training: {n} blocks, {epochs} epochs, loss {l0:.3f} -> {loss:.4f}, {t_train:.2f} s", file=sys.stderr,
          flush=True)
    ctx.attach_mlp(mlp)
    st_ln, u_ln = solve(_lib.PC_LEARNED, learned_fallback_iters=50)
    print(f"learned: {st_ln['seconds']:.3f} s, newton {st_ln['newton_steps']}, gmres {st_ln['gmres_iters']}, "
          f"fallbacks {st_ln['fallback_steps']}", file=sys.stderr, flush=True)
    st_bj32, u_bj32 = solve(_lib.PC_BLOCK_JACOBI, jvp_fp32=True)
    print(f"block-Jacobi, FP32 J.v: {st_bj32['seconds']:.3f} s, newton {st_bj32['newton_steps']}, "
          f"gmres {st_bj32['gmres_iters']}", file=sys.stderr, flush=True)
    st_sgs, u_sgs = solve(_lib.PC_BLOCK_SGS)
    print(f"block-SGS: {st_sgs['seconds']:.3f} s, newton {st_sgs['newton_steps']}, gmres {st_sgs['gmres_iters']}",
          file=sys.stderr, flush=True)
    st_s32, u_s32 = solve(_lib.PC_BLOCK_SGS32)
    print(f"block-SGS (FP32 blocks): {st_s32['seconds']:.3f} s, newton {st_s32['newton_steps']}, "
          f"gmres {st_s32['gmres_iters']}", file=sys.stderr, flush=True)
    diff = ((u_ln - u_bj).abs().max() / u_bj.abs().max()).item()
    out = {"workload": "2-D Euler DG p=3, 64x64 curved-bump channel, Mach 0.5, Roe; Newton-GMRES(60) to 1e-10",
           "block_jacobi": leg(st_bj),
           "block_jacobi_fp32_jvp": {**leg(st_bj32), "vs_fp64_jvp_wall": round(st_bj32["seconds"] / st_bj["seconds"], 4),
                                     "steady_states_rel_diff": ((u_bj32 - u_bj).abs().max() / u_bj.abs().max()).item()},
           "learned_with_fallback": leg(st_ln), "block_sgs": leg(st_sgs),
           "sgs_vs_bj_wall": round(st_sgs["seconds"] / st_bj["seconds"], 4),
           "sgs_steady_states_rel_diff": ((u_sgs - u_bj).abs().max() / u_bj.abs().max()).item(),
           "block_sgs_fp32_blocks": leg(st_s32),
           "sgs32_vs_bj_wall": round(st_s32["seconds"] / st_bj["seconds"], 4),
           "sgs32_steady_states_rel_diff": ((u_s32 - u_bj).abs().max() / u_bj.abs().max()).item(),
           "training": {"blocks": n, "epochs": epochs, "seconds": round(t_train, 3), "loss_untrained": l0,
                        "loss_trained": loss, "structure": "Kronecker rank 8, MLP 96-256-256-2176 BF16/FP32",
                        "mfma": _training_mfma(n, epochs, t_train)},
           "learned_vs_bj_wall": round(st_ln["seconds"] / st_bj["seconds"], 4),
           "steady_states_rel_diff": diff}
    if cpu:
        out["cpu_block_jacobi_estimate"] = cpu_steady_estimate(st_bj["newton_steps"], st_bj["gmres_iters"])
        print(f"cpu block-Jacobi estimate: {out['cpu_block_jacobi_estimate']['seconds_estimate']} s",
              file=sys.stderr, flush=True)
    mlp.close()
    ctx.close()
    return out


def ns_time_to_steady(torch, dev, max_newton=150):
    """Time to max|R| < 1e-10 (FP64) of the steady Navier-Stokes Taylor-Couette flow: the reference's annulus
    (r 0.5 to 1, inner cylinder rotating at omega 1 with the no_slip_all wall, outer wall adiabatic), mu 1e-3 and
    Pr 0.72 of its solution.data (Re about 250), BR2 with the Roe flux, DG p=3 on 16 x 64 curved QUADs (65,536
    DOF), from the fluid at rest, pseudo-time Newton-GMRES(60) (rtol 1e-4) with cfl_initial 1 and the ramped CFL
    capped at 300.  Legs: multicolour block SGS over the diagonal and full BR2 face blocks (FP64 and FP32-stored,
    and FP32-stored with the GMRES operator in FP32 dual numbers), and block-Jacobi, each given at most max_newton
    Newton steps."""
    import numpy as np

    from dpgsolver_amd import _lib, mesh as M

    mu, Pr = 1e-3, 0.72
    mesh = M.annulus_mesh(3, 16, 64)
    ctx = _lib.Context(2, 3, M.PDE_NAVIER_STOKES, M.FLUX_ROE, mu=mu, Pr=Pr, walls=M.taylor_couette_walls(mu, Pr))
    ctx.set_stream(torch.cuda.current_stream(dev).cuda_stream)
    ctx.upload_mesh(mesh.xyz, mesh.faces)
    u0 = M.uniform_state(mesh, np.array([1.0, 0.0, 0.0, 1.0]))
    kw = dict(cfl_initial=1.0, cfl_max=300.0, exit_tol=1e-10, max_newton=max_newton, gmres_restart=60,
              gmres_rtol=1e-4, gmres_max_iter=1000, raise_on_error=False)

    def solve(pc, **extra):
        u = torch.from_numpy(u0).to(dev)
        st = ctx.solve_implicit(u.data_ptr(), precond=pc, **{**kw, **extra})
        return {"seconds": round(st["seconds"], 4), "newton_steps": st["newton_steps"],
                "gmres_iters": st["gmres_iters"], "converged": bool(st["converged"]), "max_rhs": st["max_rhs"]}, u

    solve(_lib.PC_BLOCK_SGS, max_newton=2)      # warm-up
    sgs, u_sgs = solve(_lib.PC_BLOCK_SGS)
    print(f"navier-stokes taylor-couette, block SGS: {sgs['seconds']:.3f} s, newton {sgs['newton_steps']}, "
          f"gmres {sgs['gmres_iters']}", file=sys.stderr, flush=True)
    s32, u_s32 = solve(_lib.PC_BLOCK_SGS32)
    j32, u_j32 = solve(_lib.PC_BLOCK_SGS32, jvp_fp32=True)
    print(f"navier-stokes taylor-couette, block SGS (FP32 blocks, FP32 operator): {j32['seconds']:.3f} s, "
          f"newton {j32['newton_steps']}, gmres {j32['gmres_iters']}", file=sys.stderr, flush=True)
    bj, _ = solve(_lib.PC_BLOCK_JACOBI)
    print(f"navier-stokes taylor-couette, block-Jacobi: converged {bj['converged']}, max_rhs {bj['max_rhs']:.2e} "
          f"after {bj['newton_steps']} steps", file=sys.stderr, flush=True)
    ctx.close()
    return {"workload": "2-D Navier-Stokes BR2 (Roe), Taylor-Couette annulus r 0.5-1, mu 1e-3, Pr 0.72, omega 1, "
                        "DG p=3 on 16x64 curved QUADs (65536 DOF), from rest; Newton-GMRES(60) to 1e-10, CFL cap 300",
            "block_sgs": sgs, "block_sgs_fp32_blocks": {**s32, "steady_states_rel_diff":
                                                        ((u_s32 - u_sgs).abs().max() / u_sgs.abs().max()).item()},
            "block_sgs_fp32_blocks_fp32_jvp": {**j32, "steady_states_rel_diff":
                                               ((u_j32 - u_sgs).abs().max() / u_sgs.abs().max()).item()},
            "block_jacobi": bj}


def ns_throughput(torch, dev, steps, warmup, n=448):
    """2-D Navier-Stokes (BR2) FP64 residual on the QUAD part of the mixed-mesh workload: periodic [0,1]^2,
    n x n elements, p = 3, Roe, Pr 0.72, Re 100 on the velocity amplitude 0.28 and unit length (mu = 0.0028);
    state: the 3-D workload's formula in 2-D.  One evaluation = the gradient kernel + the residual kernel,
    replayed as one HIP graph."""
    from dpgsolver_amd import _lib, mesh as M

    p, mu, Pr = 3, 0.28 / 100.0, 0.72
    mesh = M.config4_mesh(n, p)
    ctx = _lib.Context(2, p, M.PDE_NAVIER_STOKES, M.FLUX_ROE, mu=mu, Pr=Pr)
    stream = torch.cuda.current_stream(dev)
    ctx.set_stream(stream.cuda_stream)
    ctx.upload_mesh(mesh.xyz, mesh.faces)
    u = torch.from_numpy(M.config4_state(mesh)).to(dev)
    r = torch.empty_like(u)
    g = ctx.graph_rhs(_lib.GRAPH_RHS_FP64, u.data_ptr(), r.data_ptr())
    u32 = u.float()
    r32 = torch.empty_like(u32)
    g32 = ctx.graph_rhs(_lib.GRAPH_RHS_FP32, u32.data_ptr(), r32.data_ptr())

    def timed(graph):
        for _ in range(warmup):
            graph.launch(1)
        torch.cuda.synchronize()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        graph.launch(steps)
        stop.record(stream)
        torch.cuda.synchronize()
        return start.elapsed_time(stop) / steps

    ms = timed(g)
    ms32 = timed(g32)
    err32 = ((r32.double() - r).abs().max() / r.abs().max()).item()
    rmx = torch.empty_like(u32)
    gmx = ctx.graph_rhs(_lib.GRAPH_RHS_MIXED, u.data_ptr(), rmx.data_ptr())
    msmx = timed(gmx)
    errmx = ((rmx.double() - r).abs().max() / r.abs().max()).item()
    gmx.close()
    # input rounding alone: the FP64 residual of the FP32-rounded state
    ur = u32.double()
    rr = torch.empty_like(u)
    ctx.compute_rhs(ur.data_ptr(), rr.data_ptr())
    err_in = ((rr - r).abs().max() / r.abs().max()).item()
    del ur, rr
    E, Np, nv, d, Nfc = mesh.E, 16, 4, 2, 4
    ndof = E * Np * nv
    # compulsory traffic: state in, residual out, FP64 geometry (metrics, det J, n J_f of the owned faces)
    alg = E * (2 * nv * Np * 8 + Np * (d * d + 1) * 8 + d * Nfc * d * 8)
    # plus the two-kernel design's gradient round trip (G at the GL nodes and the face traces, written, read)
    grad_rt = E * 2 * (d * nv * Np * 8 + 2 * d * d * nv * Nfc * 8)
    alg32 = E * (2 * nv * Np * 4 + Np * (d * d + 1) * 4 + d * Nfc * d * 4)
    g.close()
    g32.close()
    ctx.close()
    return {"workload": f"2-D Navier-Stokes BR2 p={p}, {n}x{n} periodic QUAD ({E} elements), Roe, Re 100, Pr 0.72",
            "fp32": {"value": round(ndof / (ms32 * 1e-3) / 1e9, 3), "ms_per_step": round(ms32, 5),
                     "rel_err_vs_fp64": err32, "input_rounding_rel_err": err_in,
                     "roofline": {"bound": "hbm", "achieved": round(alg32 / (ms32 * 1e-3) / 1e9, 1),
                                  "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": round(alg32 / (ms32 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                                  "traffic": _traffic_from_profiles("ns_fp32", "ns_quad_448_p3_traffic.md")}},
            "mixed": {"value": round(ndof / (msmx * 1e-3) / 1e9, 3), "ms_per_step": round(msmx, 5),
                      "rel_err_vs_fp64": errmx},
            "dtype": "f64", "value": round(ndof / (ms * 1e-3) / 1e9, 3), "unit": "GDOF/s", "ms_per_step": round(ms, 5),
            "roofline": {"bound": "hbm", "achieved": round(alg / (ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                         "traffic": _traffic_from_profiles("ns_fp64", "ns_quad_448_p3_traffic.md")},
            "gradient_round_trip_bytes": grad_rt}


def hp_ns_throughput(torch, dev, steps, warmup):
    """2-D Navier-Stokes (BR2) residual on the variable-order workload: the 500 x 300 periodic QUAD base grid
    of [0,1]^2 with orders p ~ U{1..5} per element (seeded; the TRI split is not built), Roe, Pr 0.72, Re 100
    (mu = 0.0028).  One evaluation = every order group's gradient kernel, the face kernel between orders and
    every group's residual kernel, replayed as one HIP graph (FP64, FP32 and the mixed form)."""
    import numpy as np
    from dpgsolver_amd import _lib, mesh as M

    mu, Pr = 0.28 / 100.0, 0.72
    mesh = M.config4_hp_mesh()
    ctx = _lib.Context(2, 5, M.PDE_NAVIER_STOKES, M.FLUX_ROE, mu=mu, Pr=Pr)
    stream = torch.cuda.current_stream(dev)
    ctx.set_stream(stream.cuda_stream)
    ctx.upload_mesh_hp(mesh)
    u = torch.from_numpy(M.config4_hp_state(mesh)).to(dev)
    r = torch.empty_like(u)
    u32 = u.float()
    r32 = torch.empty_like(u32)
    g = ctx.graph_rhs(_lib.GRAPH_RHS_FP64, u.data_ptr(), r.data_ptr())
    g32 = ctx.graph_rhs(_lib.GRAPH_RHS_FP32, u32.data_ptr(), r32.data_ptr())

    def timed(graph):
        for _ in range(warmup):
            graph.launch(1)
        torch.cuda.synchronize()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        graph.launch(steps)
        stop.record(stream)
        torch.cuda.synchronize()
        return start.elapsed_time(stop) / steps

    ms = timed(g)
    ms32 = timed(g32)
    err32 = ((r32.double() - r).abs().max() / r.abs().max()).item()
    rmx = torch.empty_like(u32)
    gmx = ctx.graph_rhs(_lib.GRAPH_RHS_MIXED, u.data_ptr(), rmx.data_ptr())
    msmx = timed(gmx)
    errmx = ((rmx.double() - r).abs().max() / r.abs().max()).item()
    gmx.close()
    ur = u32.double()
    rr = torch.empty_like(u)
    ctx.compute_rhs(ur.data_ptr(), rr.data_ptr())
    torch.cuda.synchronize()
    err_in = ((rr - r).abs().max() / r.abs().max()).item()
    del ur, rr
    nv, d = 4, 2
    Np = (mesh.p.astype(np.int64) + 1) ** 2
    N = mesh.p.astype(np.int64) + 1
    ndof = int((Np * nv).sum())
    # compulsory traffic per element: state in, residual out, geometry (metrics, det J, n J_f of its d owned faces)
    alg = int((2 * nv * Np * 8 + Np * (d * d + 1) * 8 + d * N * d * 8).sum())
    alg32 = int((2 * nv * Np * 4 + Np * (d * d + 1) * 4 + d * N * d * 4).sum())
    fp, fo = mesh.p[mesh.faces[:, 0]], mesh.p[mesh.faces[:, 2]]
    g.close()
    g32.close()
    ctx.close()
    return {"workload": f"2-D Navier-Stokes BR2, 500x300 periodic QUAD ({mesh.E} elements), p ~ U{{1..5}} per "
                        f"element, Roe, Re 100, Pr 0.72",
            "elements_per_order": np.bincount(mesh.p, minlength=6)[1:].tolist(),
            "faces_between_orders": int((fp != fo).sum()), "faces": int(mesh.F), "ndof": ndof,
            "dtype": "f64", "value": round(ndof / (ms * 1e-3) / 1e9, 3), "unit": "GDOF/s",
            "ms_per_step": round(ms, 5),
            "roofline": {"bound": "hbm", "achieved": round(alg / (ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                         "traffic": _traffic_from_profiles("hp_ns_fp64", "variable_order_500x300_traffic.md")},
            "fp32": {"value": round(ndof / (ms32 * 1e-3) / 1e9, 3), "ms_per_step": round(ms32, 5),
                     "rel_err_vs_fp64": err32, "input_rounding_rel_err": err_in,
                     "roofline": {"bound": "hbm", "achieved": round(alg32 / (ms32 * 1e-3) / 1e9, 1),
                                  "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": round(alg32 / (ms32 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}},
            "mixed": {"value": round(ndof / (msmx * 1e-3) / 1e9, 3), "ms_per_step": round(msmx, 5),
                      "rel_err_vs_fp64": errmx}}


def hp_euler_throughput(torch, dev, steps, warmup):
    """2-D Euler residual on the same variable-order grid (500 x 300 periodic QUAD, p ~ U{1..5}, Roe): the face
    kernel between orders and every order group's residual kernel as one HIP graph, FP64, FP32 and mixed."""
    from dpgsolver_amd import _lib, mesh as M

    mesh = M.config4_hp_mesh()
    ctx = _lib.Context(2, 5, M.PDE_EULER, M.FLUX_ROE)
    stream = torch.cuda.current_stream(dev)
    ctx.set_stream(stream.cuda_stream)
    ctx.upload_mesh_hp(mesh)
    u = torch.from_numpy(M.config4_hp_state(mesh)).to(dev)
    u32 = u.float()
    outs = {_lib.GRAPH_RHS_FP64: torch.empty_like(u), _lib.GRAPH_RHS_FP32: torch.empty_like(u32),
            _lib.GRAPH_RHS_MIXED: torch.empty_like(u32)}
    ins = {_lib.GRAPH_RHS_FP64: u, _lib.GRAPH_RHS_FP32: u32, _lib.GRAPH_RHS_MIXED: u}
    ms = {}
    for kind, out in outs.items():
        g = ctx.graph_rhs(kind, ins[kind].data_ptr(), out.data_ptr())
        for _ in range(warmup):
            g.launch(1)
        torch.cuda.synchronize()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        g.launch(steps)
        stop.record(stream)
        torch.cuda.synchronize()
        ms[kind] = start.elapsed_time(stop) / steps
        g.close()
    r = outs[_lib.GRAPH_RHS_FP64]
    err = {k: ((outs[k].double() - r).abs().max() / r.abs().max()).item()
           for k in (_lib.GRAPH_RHS_FP32, _lib.GRAPH_RHS_MIXED)}
    ndof = u.numel()
    ctx.close()
    leg = lambda k: {"value": round(ndof / (ms[k] * 1e-3) / 1e9, 3), "ms_per_step": round(ms[k], 5)}
    return {"workload": "2-D Euler, 500x300 periodic QUAD, p ~ U{1..5} per element, Roe", "ndof": ndof,
            "fp64": leg(_lib.GRAPH_RHS_FP64),
            "fp32": {**leg(_lib.GRAPH_RHS_FP32), "rel_err_vs_fp64": err[_lib.GRAPH_RHS_FP32]},
            "mixed": {**leg(_lib.GRAPH_RHS_MIXED), "rel_err_vs_fp64": err[_lib.GRAPH_RHS_MIXED]}}


def mixed_throughput(torch, dev, steps, warmup, pde):
    """2-D FP64 and FP32 residuals on configuration 4's mesh: the 500 x 300 periodic base grid of [0,1]^2 with a seeded third
    of its cells cut into two TRIs (99,980 TRI + 100,010 QUAD), p ~ U{1..5} per element (seeded), Roe; the state is
    the 3-D workload's formula in 2-D.  Euler, or Navier-Stokes (BR2) with Re 100, Pr 0.72 (mu = 0.0028), the
    configuration's own equations.  One evaluation = (Navier-Stokes: every group's gradient kernel,) the face
    kernels (between QUAD orders, and every face with a TRI side) and every (type, p) group's residual kernel,
    replayed as one HIP graph."""
    import numpy as np
    from dpgsolver_amd import _lib, mesh as M

    ns = pde == "navier_stokes"
    mesh = M.config4_mixed_mesh()
    if ns:
        ctx = _lib.Context(2, 5, M.PDE_NAVIER_STOKES, M.FLUX_ROE, mu=0.28 / 100.0, Pr=0.72)
    else:
        ctx = _lib.Context(2, 5, M.PDE_EULER, M.FLUX_ROE)
    stream = torch.cuda.current_stream(dev)
    ctx.set_stream(stream.cuda_stream)
    ctx.upload_mesh_mixed(mesh)
    u = torch.from_numpy(M.config4_mixed_state(mesh)).to(dev)
    r = torch.empty_like(u)
    u32 = u.float()
    r32 = torch.empty_like(u32)

    def timed(kind, ui, ri):
        g = ctx.graph_rhs(kind, ui.data_ptr(), ri.data_ptr())
        for _ in range(warmup):
            g.launch(1)
        torch.cuda.synchronize()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        g.launch(steps)
        stop.record(stream)
        torch.cuda.synchronize()
        g.close()
        return start.elapsed_time(stop) / steps

    ms = timed(_lib.GRAPH_RHS_FP64, u, r)
    ms32 = timed(_lib.GRAPH_RHS_FP32, u32, r32)
    # FP32 error against FP64, and the part that rounding the state alone to FP32 causes
    rr = torch.empty_like(u)
    u64r = u32.double()
    ctx.compute_rhs(u64r.data_ptr(), rr.data_ptr())
    torch.cuda.synchronize()
    scale = r.abs().max().item()
    err32 = (r32.double() - r).abs().max().item() / scale
    err_in = (rr - r).abs().max().item() / scale
    del rr, u64r
    ctx.close()
    tri = mesh.kind == 1
    p = mesh.p.astype(np.int64)
    Np = np.where(tri, (p + 1) * (p + 2) // 2, (p + 1) ** 2)
    Nc = np.where(tri, np.array([0, 3, 6, 12, 16, 25])[p], (p + 1) ** 2)
    # compulsory bytes: state in and residual out, the volume geometry at the cubature (QUAD: 4 metrics and det J;
    # TRI: 4 metrics), n J_f at the GL points of every face (2 values); Navier-Stokes on these straight (affine)
    # TRIs: also 1 / det J per element (their M^-1 and lift traces are the reference ones scaled by it)
    geo = np.where(tri, 4 * Nc, 5 * Nc) * 8
    if ns:
        geo = geo + np.where(tri, 1, 0) * 8
    fmax = np.maximum(mesh.p[mesh.faces[:, 0]], np.where(mesh.faces[:, 2] >= 0, mesh.p[mesh.faces[:, 2]], 0)) + 1
    alg = int((2 * 4 * Np * 8 + geo).sum() + (fmax * 2 * 8).sum())
    ndof = u.numel()
    eq = "Navier-Stokes BR2 (Re 100, Pr 0.72)" if ns else "Euler"
    return {"workload": f"2-D {eq}, configuration 4's mesh: 500x300 periodic grid, a seeded third of the cells "
                        "cut into two TRIs, p ~ U{1..5} per element, Roe", "elements": int(mesh.E),
            "tri_elements": int(tri.sum()), "faces": int(mesh.F), "ndof": ndof, "dtype": "f64",
            "value": round(ndof / (ms * 1e-3) / 1e9, 3), "unit": "GDOF/s", "ms_per_step": round(ms, 5),
            "roofline": {"bound": "hbm", "achieved": round(alg / (ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                         "traffic": _traffic_from_profiles("mixed_ns_fp64" if ns else "mixed_euler_fp64",
                                                           "mixed_tri_quad_500x300_traffic.md")},
            # FP32 state, geometry shadows and operators: the same values in half the bytes
            "fp32": {"value": round(ndof / (ms32 * 1e-3) / 1e9, 3), "ms_per_step": round(ms32, 5),
                     "rel_err_vs_fp64": err32, "input_rounding_rel_err": err_in,
                     "roofline": {"bound": "hbm", "achieved": round(alg / 2 / (ms32 * 1e-3) / 1e9, 1),
                                  "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": round(alg / 2 / (ms32 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "traffic": None}}}


def config2_throughput(torch, dev, steps, warmup):
    """2-D Euler, DG p=3 on the 128 x 128 periodic QUAD mesh of [0,10]^2, isentropic vortex (strength 5, radius 1,
    Mach 0.5 along x) plus a seeded 1e-3 perturbation, Roe-Pike: the FP64 and mixed residuals as graph replays,
    and one whole SSP-RK3 step (three M^-1 R evaluations, stage updates, positivity limiter; time_step_ssp_rk_33,
    solve_explicit.c:252-301) replayed as one HIP graph."""
    from dpgsolver_amd import _lib, mesh as M

    mesh = M.config2_mesh(128, 3)
    ctx = _lib.Context(2, 3, M.PDE_EULER, M.FLUX_ROE)
    stream = torch.cuda.current_stream(dev)
    ctx.set_stream(stream.cuda_stream)
    ctx.upload_mesh(mesh.xyz, mesh.faces)
    u = torch.from_numpy(M.config2_state(mesh)).to(dev)
    r = torch.empty_like(u)
    rmx = torch.empty(u.numel(), dtype=torch.float32, device=dev)

    def timed(graph, k):
        for _ in range(warmup):
            graph.launch(1)
        torch.cuda.synchronize()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        graph.launch(k)
        stop.record(stream)
        torch.cuda.synchronize()
        return start.elapsed_time(stop) / k

    g64 = ctx.graph_rhs(_lib.GRAPH_RHS_FP64, u.data_ptr(), r.data_ptr())
    gmx = ctx.graph_rhs(_lib.GRAPH_RHS_MIXED, u.data_ptr(), rmx.data_ptr())
    ms64, msmx = timed(g64, steps), timed(gmx, steps)
    err = ((rmx.double() - r).abs().max() / r.abs().max()).item()
    g64.close()
    gmx.close()
    # SSP-RK3 steps: dt = 0.2 h / ((2p+1) max(|V| + c)) on the state, the steps replayed on a copy
    h = 10.0 / 128
    dt = 0.2 * h / (7 * 2.0)
    us = u.clone()
    grk = ctx.graph_rk3(us.data_ptr(), dt, positivity=True)
    ms_rk = timed(grk, steps)
    mx = grk.sync()
    grk.close()
    ndof = u.numel()
    ctx.close()
    return {"workload": "2-D Euler DG p=3, 128x128 periodic QUAD on [0,10]^2, isentropic vortex + 1e-3 noise, Roe",
            "ndof": ndof,
            "fp64": {"value": round(ndof / (ms64 * 1e-3) / 1e9, 3), "unit": "GDOF/s", "ms_per_step": round(ms64, 5)},
            "mixed": {"value": round(ndof / (msmx * 1e-3) / 1e9, 3), "unit": "GDOF/s", "ms_per_step": round(msmx, 5),
                      "rel_err_vs_fp64": err},
            "ssp_rk3": {"steps_per_s": round(1e3 / ms_rk, 1), "ms_per_step": round(ms_rk, 5), "dt": dt,
                        "max_abs_minv_r_last_stage": mx,
                        "note": "one HIP graph replay = one SSP-RK3 step with the positivity limiter, FP64"}}


def config1_rk3(torch, dev, cpu=True):
    """2-D linear advection, DG p=2 on the 16 x 16 periodic QUAD mesh of [-1,1]^2, b = (1, 0.5), upwind flux,
    u0 = sin(pi x) sin(pi y), SSP-RK3 with dt = 0.2 h / ((2p+1) |b|_inf), 100 steps, FP64: the GPU steps as HIP
    graph replays; beside them, as the CPU baseline and the check, the oracle's same 100 steps on one core."""
    import numpy as np
    from dpgsolver_amd import _lib, mesh as M

    mesh = M.config1_mesh(16, 2)
    ctx = _lib.Context(2, 2, M.PDE_ADVECTION, M.FLUX_UPWIND, b_adv=(1.0, 0.5, 0.0))
    stream = torch.cuda.current_stream(dev)
    ctx.set_stream(stream.cuda_stream)
    ctx.upload_mesh(mesh.xyz, mesh.faces)
    u0 = M.config1_state(mesh)
    dt, nsteps = 0.2 * 0.125 / 5.0, 100
    u = torch.from_numpy(u0).to(dev)
    g = ctx.graph_rk3(u.data_ptr(), dt, positivity=False)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    g.launch(nsteps)
    stop.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop)
    g.close()
    ctx.close()
    u_gpu = u.cpu().numpy()
    T = nsteps * dt
    x = M.nodes_xyz(mesh)
    exact = (np.sin(np.pi * (x[..., 0] - T)) * np.sin(np.pi * (x[..., 1] - 0.5 * T))).reshape(-1)
    out = {"workload": "2-D advection DG p=2, 16x16 periodic QUAD, upwind, SSP-RK3 100 steps (FP64)",
           "gpu_ms_total": round(ms, 4), "gpu_steps_per_s": round(nsteps / (ms * 1e-3), 1),
           "max_abs_error_vs_exact_translation": float(np.abs(u_gpu - exact).max())}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        orc = pyoracle.Oracle(mesh, M.PDE_ADVECTION, M.FLUX_UPWIND, b_adv=(1.0, 0.5, 0.0))
        uc = u0.copy()
        t0 = time.perf_counter()
        for _ in range(nsteps):
            uc, _, _ = orc.ssp_rk3_step(uc, dt, positivity=False)
        t_cpu = time.perf_counter() - t0
        out["cpu_oracle_1_core"] = {"seconds": round(t_cpu, 4), "steps_per_s": round(nsteps / t_cpu, 1),
                                    "kind": "port", "cores": 1}
        out["final_state_rel_diff_vs_oracle"] = float(np.abs(u_gpu - uc).max() / np.abs(uc).max())
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=64, help="elements per direction (64 = the named workload)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-implicit", action="store_true", help="skip the time-to-1e-10 steady-solve legs")
    ap.add_argument("--implicit-only", action="store_true",
                    help="run only the time-to-1e-10 legs and print their JSON (profiling aid)")
    ap.add_argument("--ns-only", action="store_true",
                    help="run only the Navier-Stokes residual leg and print its JSON (profiling aid)")
    ap.add_argument("--hp-only", action="store_true",
                    help="run only the variable-order Navier-Stokes leg and print its JSON (profiling aid)")
    ap.add_argument("--mixed-only", action="store_true",
                    help="run only the mixed TRI/QUAD legs and print their JSON (profiling aid)")
    ap.add_argument("--cpu-steady", action="store_true",
                    help="run only the measured CPU time-to-1e-10 (host cores) beside the GPU solve and print its JSON")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one 1 GiB read + 1 GiB write elementwise kernel (calibrates rocprofv3 FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "sharding is a design, not a measured path.", file=sys.stderr)
        sys.exit(2)

    if args.cpu_steady:
        print(json.dumps({"cpu_time_to_1e-10": cpu_steady_measured()}), flush=True)
        return

    import numpy as np
    import torch

    from dpgsolver_amd import _lib, mesh as M

    dev = torch.device("cuda:0")
    if args.implicit_only:
        print(json.dumps({"time_to_1e-10": time_to_steady(torch, dev, cpu=not args.no_cpu),
                          "navier_stokes_time_to_1e-10": ns_time_to_steady(torch, dev)}), flush=True)
        return
    if args.ns_only:
        print(json.dumps({"navier_stokes": ns_throughput(torch, dev, args.steps, args.warmup)}), flush=True)
        return
    if args.mixed_only:
        print(json.dumps({"euler_mixed_tri_quad": mixed_throughput(torch, dev, args.steps, args.warmup, "euler"),
                          "navier_stokes_mixed_tri_quad": mixed_throughput(torch, dev, args.steps, args.warmup,
                                                                           "navier_stokes")}), flush=True)
        return
    if args.hp_only:
        print(json.dumps({"navier_stokes_variable_order": hp_ns_throughput(torch, dev, args.steps, args.warmup),
                          "euler_variable_order": hp_euler_throughput(torch, dev, args.steps, args.warmup)}),
              flush=True)
        return
    n, p = args.n, 4
    t_setup = time.perf_counter()
    mesh = M.config3_mesh(n, p)
    ctx = _lib.Context(3, p, M.PDE_EULER, M.FLUX_LF)
    stream = torch.cuda.current_stream(dev)
    ctx.set_stream(stream.cuda_stream)
    ctx.upload_mesh(mesh.xyz, mesh.faces)
    xyz_dev = torch.from_numpy(mesh.xyz).to(dev)
    u64 = _state_on_device(torch, xyz_dev, M.SEED)
    del xyz_dev
    ndof = u64.numel()
    u32 = torch.empty(ndof, dtype=torch.float32, device=dev)
    ctx.convert_f64_f32(u64.data_ptr(), u32.data_ptr(), ndof)
    r64 = torch.empty_like(u64)
    r32 = torch.empty_like(u32)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t_setup
    print(f"setup: mesh {mesh.E} elements, {ndof} DOF, {setup_s:.1f} s", file=sys.stderr, flush=True)

    def timed(fn, k):
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        start.record(stream)
        for _ in range(k):
            fn()
        stop.record(stream)
        torch.cuda.synchronize()
        return start.elapsed_time(stop) / k   # ms per launch

    # launch shapes from the autotuner (one workgroup per element block, or a persistent grid), then each
    # evaluation is one replay of a HIP graph holding the residual launch (dgx_graph_create_rhs)
    rmx = torch.empty_like(u32)
    tuned = {}
    for name, kind, ui, ri in (("fp32", _lib.GRAPH_RHS_FP32, u32, r32), ("fp64", _lib.GRAPH_RHS_FP64, u64, r64),
                               ("mixed", _lib.GRAPH_RHS_MIXED, u64, rmx)):
        grid, ms = ctx.autotune_rhs(kind, ui.data_ptr(), ri.data_ptr())
        tuned[name] = {"grid": grid, "ms": round(ms, 5)}
    print(f"autotuned grids: {tuned}", file=sys.stderr, flush=True)
    g32 = ctx.graph_rhs(_lib.GRAPH_RHS_FP32, u32.data_ptr(), r32.data_ptr())
    g64 = ctx.graph_rhs(_lib.GRAPH_RHS_FP64, u64.data_ptr(), r64.data_ptr())
    gmx = ctx.graph_rhs(_lib.GRAPH_RHS_MIXED, u64.data_ptr(), rmx.data_ptr())
    f32 = lambda: g32.launch(1)
    f64 = lambda: g64.launch(1)
    fmx = lambda: gmx.launch(1)
    for _ in range(args.warmup):
        f32()
        f64()
        fmx()
    torch.cuda.synchronize()
    ms32_eager = timed(lambda: ctx.compute_rhs_f32(u32.data_ptr(), r32.data_ptr()), args.steps)
    ms32 = timed(f32, args.steps)
    print(f"fp32 residual: {ms32:.4f} ms", file=sys.stderr, flush=True)
    ms64 = timed(f64, args.steps)
    print(f"fp64 residual: {ms64:.4f} ms", file=sys.stderr, flush=True)
    msmx = timed(fmx, args.steps)
    print(f"mixed residual: {msmx:.4f} ms", file=sys.stderr, flush=True)

    err32 = ((r32.double() - r64).abs().max() / r64.abs().max()).item()
    errmx = ((rmx.double() - r64).abs().max() / r64.abs().max()).item()
    # input rounding alone: FP64 residual of the FP32-rounded state
    u64r = u32.double()
    r64r = torch.empty_like(u64)
    ctx.compute_rhs(u64r.data_ptr(), r64r.data_ptr())
    err_in = ((r64r - r64).abs().max() / r64.abs().max()).item()
    del u64r, r64r
    # size-independent property at full size: free-stream preservation on the curved mesh
    prim = torch.tensor([1.0, 0.3, -0.2, 0.1, 1.0], dtype=torch.float64, device=dev)
    cons0 = torch.stack([prim[0], prim[0] * prim[1], prim[0] * prim[2], prim[0] * prim[3],
                         prim[4] / M.GM1 + 0.5 * prim[0] * (prim[1:4] ** 2).sum()])
    ufs = cons0.view(1, 5, 1).expand(mesh.E, 5, 125).contiguous().view(-1)
    ctx.compute_rhs(ufs.data_ptr(), r64.data_ptr())
    torch.cuda.synchronize()
    free_stream = r64.abs().max().item()
    del ufs
    if args.copy_probe:
        a = torch.ones(1 << 27, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        torch.mul(a, 2.0, out=b)   # one elementwise kernel: reads 1 GiB, writes 1 GiB
        torch.cuda.synchronize()
        del a, b

    contraction_kind = "sum-factorised 1-D sweeps on the VALU"
    E, Np, nv, d, Nfc = mesh.E, 125, 5, 3, 25
    bytes32 = E * (2 * nv * Np * 4 + Np * d * d * 4 + d * Nfc * d * 4)
    bytes64 = E * (2 * nv * Np * 8 + Np * d * d * 8 + d * Nfc * d * 8)
    bytesmx = E * (nv * Np * 8 + nv * Np * 4 + Np * d * d * 4 + d * Nfc * d * 4)
    gdof32 = ndof / (ms32 * 1e-3) / 1e9
    gdof64 = ndof / (ms64 * 1e-3) / 1e9
    ach32 = bytes32 / (ms32 * 1e-3) / 1e9
    ach64 = bytes64 / (ms64 * 1e-3) / 1e9
    achmx = bytesmx / (msmx * 1e-3) / 1e9
    gdofmx = ndof / (msmx * 1e-3) / 1e9

    ns = ns_throughput(torch, dev, args.steps, args.warmup)
    print(f"navier-stokes residual: {ns['ms_per_step']:.4f} ms, {ns['value']} GDOF/s", file=sys.stderr, flush=True)
    hp = hp_ns_throughput(torch, dev, args.steps, args.warmup)
    print(f"variable-order navier-stokes residual: {hp['ms_per_step']:.4f} ms, {hp['value']} GDOF/s",
          file=sys.stderr, flush=True)
    hpe = hp_euler_throughput(torch, dev, args.steps, args.warmup)
    print(f"variable-order euler residual: {hpe['fp64']['ms_per_step']:.4f} ms, {hpe['fp64']['value']} GDOF/s",
          file=sys.stderr, flush=True)
    mx4 = mixed_throughput(torch, dev, args.steps, args.warmup, "euler")
    print(f"mixed TRI/QUAD euler residual: {mx4['ms_per_step']:.4f} ms, {mx4['value']} GDOF/s (FP32 {mx4['fp32']['ms_per_step']:.4f} ms)", file=sys.stderr,
          flush=True)
    mx4ns = mixed_throughput(torch, dev, args.steps, args.warmup, "navier_stokes")
    print(f"mixed TRI/QUAD navier-stokes residual: {mx4ns['ms_per_step']:.4f} ms, {mx4ns['value']} GDOF/s (FP32 {mx4ns['fp32']['ms_per_step']:.4f} ms)",
          file=sys.stderr, flush=True)
    c2 = config2_throughput(torch, dev, args.steps, args.warmup)
    print(f"config-2 euler residual: fp64 {c2['fp64']['value']} GDOF/s, mixed {c2['mixed']['value']} GDOF/s, "
          f"rk3 {c2['ssp_rk3']['ms_per_step']} ms/step", file=sys.stderr, flush=True)
    c1 = config1_rk3(torch, dev, cpu=not args.no_cpu)
    print(f"config-1 rk3: {c1['gpu_steps_per_s']} steps/s", file=sys.stderr, flush=True)
    cpu = None if args.no_cpu else cpu_baseline()
    implicit = None if args.no_implicit else time_to_steady(torch, dev, cpu=not args.no_cpu)
    ns_steady = None if args.no_implicit else ns_time_to_steady(torch, dev)

    out = {
        "metric": "GDOF/s residual eval (FP64 vs FP32-MFMA); time-to-1e-10 steady residual",
        "value": round(gdofmx, 3),
        "unit": "GDOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(msmx, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 contractions, fp64 state and pointwise fluxes",
        "rel_err_vs_fp64": errmx,
        "data": "synthetic",
        "config": {"workload": f"3-D Euler DG p=4, {n}^3 curved periodic HEX, Lax-Friedrichs, residual eval",
                   "elements": E, "dof": ndof, "p": p, "num_flux": "Lax-Friedrichs",
                   "contractions": contraction_kind},
        "roofline": {"bound": "hbm", "achieved": round(achmx, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achmx / HBM_PEAK_GBS, 4), "traffic": _traffic_from_profiles("rhs_mixed")},
        "fp64": {"value": round(gdof64, 3), "ms_per_step": round(ms64, 5),
                 "roofline": {"bound": "hbm", "achieved": round(ach64, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(ach64 / HBM_PEAK_GBS, 4),
                              "traffic": _traffic_from_profiles("rhs_fp64")}},
        "fp32": {"value": round(gdof32, 3), "ms_per_step": round(ms32, 5), "rel_err_vs_fp64": err32,
                 "input_rounding_rel_err": err_in,
                 "note": "FP32 state: rounding the input alone exceeds 1e-6, so this is not the inner residual",
                 "roofline": {"bound": "hbm", "achieved": round(ach32, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(ach32 / HBM_PEAK_GBS, 4), "traffic": _traffic_from_profiles("rhs_fp32")}},
        "fp32_ms_per_step_eager_launch": round(ms32_eager, 5),
        "launch_grids": tuned,
        "fp64_free_stream_max_abs": free_stream,
        "euler_mixed_tri_quad": mx4,
        "navier_stokes_mixed_tri_quad": mx4ns,
        "euler_vortex_128": c2,
        "advection_rk3_16": c1,
        "navier_stokes": ns,
        "navier_stokes_variable_order": hp,
        "euler_variable_order": hpe,
        "cpu_baseline": cpu,
        "time_to_1e-10": implicit,
        "navier_stokes_time_to_1e-10": ns_steady,
    }
    print(json.dumps(out), flush=True)
    for g in (g32, g64, gmx):
        g.close()
    ctx.close()


if __name__ == "__main__":
    main()
