"""ctypes binding of libdgx.so (the HIP C-ABI declared in include/dgx.h).

Pointers passed to the ``*_dev`` methods are HIP device addresses (e.g. ``tensor.data_ptr()``
of a torch tensor on ``cuda:0``).  The library is required: importing this module without a
built ``libdgx.so`` raises, and there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdgx.so")

DGX_OK = 0
_ERRORS = {-1: "DGX_ERR_ARG", -2: "DGX_ERR_HIP", -3: "DGX_ERR_UNSUPPORTED", -4: "DGX_ERR_POSITIVITY",
           -5: "DGX_ERR_NONFINITE", -6: "DGX_ERR_NOT_CONVERGED", -7: "DGX_ERR_NO_MESH"}

# every symbol include/dgx.h declares
EXPORTS = [
    "dgx_face_permutation", "dgx_reference_nodes", "dgx_tri_rule", "dgx_upload_mesh_mixed", "dgx_version", "dgx_device_count", "dgx_create", "dgx_destroy", "dgx_set_stream", "dgx_synchronize",
    "dgx_sizes", "dgx_upload_mesh", "dgx_upload_mesh_hp", "dgx_get_geometry", "dgx_compute_rhs", "dgx_compute_volume_rhs", "dgx_compute_face_rhs", "dgx_compute_rhs_f32", "dgx_compute_rhs_mixed",
    "dgx_compute_rhs_explicit", "dgx_apply_minv", "dgx_enforce_positivity", "dgx_ssp_rk3_step", "dgx_ssp_rk3_step_noalloc",
    "dgx_graph_create_rhs", "dgx_graph_create_rk3", "dgx_graph_launch", "dgx_graph_sync", "dgx_graph_destroy",
    "dgx_set_rhs_grid", "dgx_autotune_rhs", "dgx_max_abs",
    "dgx_convert_f64_f32", "dgx_convert_f32_f64", "dgx_assemble_diag_blocks", "dgx_jacobian_vector",
    "dgx_assemble_face_blocks", "dgx_jacobian_bsr_pattern", "dgx_assemble_jacobian_bsr", "dgx_newton_operator", "dgx_block_lu_factor", "dgx_block_lu_solve", "dgx_block_inverse", "dgx_block_inv_apply", "dgx_cfl_dt", "dgx_mass_sub", "dgx_precond_apply", "dgx_solve_implicit",
    "dgx_mesh_read_gmsh", "dgx_mesh_free", "dgx_mesh_read_gmsh_mixed", "dgx_mesh_mixed_free",
    "dgx_restart_write_gmsh", "dgx_restart_read_solution",
    "dgx_mlp_create", "dgx_mlp_destroy", "dgx_mlp_sizes", "dgx_mlp_get_params", "dgx_mlp_set_params",
    "dgx_mlp_features", "dgx_mlp_loss_grad", "dgx_mlp_train", "dgx_mlp_predict", "dgx_attach_mlp",
]

PC_NONE, PC_BLOCK_JACOBI, PC_LEARNED, PC_BLOCK_SGS, PC_BLOCK_SGS32 = 0, 1, 2, 3, 4
HISTORY_MAX = 256


class DgxError(RuntimeError):
    def __init__(self, fn: str, code: int):
        super().__init__(f"{fn} failed: {_ERRORS.get(code, code)}")
        self.code = code


class DgxConfig(C.Structure):
    _fields_ = [("d", C.c_int), ("p", C.c_int), ("pde", C.c_int), ("num_flux", C.c_int),
                ("b_adv", C.c_double * 3), ("bc_prim", C.c_double * 5), ("mu", C.c_double), ("Pr", C.c_double),
                ("wall_nf_E", C.c_double), ("wall_omega", C.c_double), ("wall_rho", C.c_double), ("wall_E", C.c_double),
                ("viscosity_type", C.c_int), ("r_s", C.c_double)]


class DgxHostMesh(C.Structure):
    _fields_ = [("d", C.c_int), ("p", C.c_int), ("E", C.c_int), ("F", C.c_int),
                ("xyz", C.POINTER(C.c_double)), ("faces", C.POINTER(C.c_int)), ("v_to_v", C.POINTER(C.c_int)),
                ("v_to_lf", C.POINTER(C.c_int)), ("v_to_lf_wp", C.POINTER(C.c_int)),
                ("n_unsupported_bc", C.c_int), ("n_inverted", C.c_int)]


class DgxHostMeshMixed(C.Structure):
    _fields_ = [("E", C.c_int), ("F", C.c_int), ("p", C.c_int), ("type", C.POINTER(C.c_int)),
                ("xyz", C.POINTER(C.c_double)), ("faces", C.POINTER(C.c_int)), ("v_to_v", C.POINTER(C.c_int)),
                ("v_to_lf", C.POINTER(C.c_int)), ("v_to_lf_wp", C.POINTER(C.c_int)),
                ("n_unsupported_bc", C.c_int), ("n_inverted", C.c_int)]


class DgxMesh(C.Structure):
    _fields_ = [("E", C.c_int), ("F", C.c_int), ("xyz", C.POINTER(C.c_double)), ("faces", C.POINTER(C.c_int))]


class DgxMeshMixed(C.Structure):
    _fields_ = [("E", C.c_int), ("F", C.c_int), ("type", C.POINTER(C.c_int)), ("p", C.POINTER(C.c_int)),
                ("xyz", C.POINTER(C.c_double)), ("faces", C.POINTER(C.c_int))]


class DgxMeshHp(C.Structure):
    _fields_ = [("E", C.c_int), ("F", C.c_int), ("p", C.POINTER(C.c_int)), ("xyz", C.POINTER(C.c_double)),
                ("faces", C.POINTER(C.c_int))]


class ImplicitParams(C.Structure):
    _fields_ = [("cfl_initial", C.c_double), ("exit_tol", C.c_double), ("exit_ratio", C.c_double),
                ("max_newton", C.c_int), ("gmres_restart", C.c_int), ("gmres_rtol", C.c_double),
                ("gmres_max_iter", C.c_int), ("precond", C.c_int), ("cfl_max", C.c_double),
                ("learned_fallback_iters", C.c_int), ("capture_blocks", C.c_void_p),
                ("capture_features", C.c_void_p), ("capture_sigma", C.c_void_p), ("capture_capacity", C.c_longlong),
                ("jvp_fp32", C.c_int)]


class ImplicitStats(C.Structure):
    _fields_ = [("newton_steps", C.c_int), ("gmres_iters", C.c_int), ("gmres_not_converged", C.c_int),
                ("converged", C.c_int), ("max_rhs0", C.c_double), ("max_rhs", C.c_double),
                ("seconds", C.c_double), ("captured", C.c_longlong), ("fallback_steps", C.c_int),
                ("gmres_per_step", C.c_int * HISTORY_MAX), ("history", C.c_double * HISTORY_MAX)]


_lib = None
vp = C.c_void_p
dp = C.POINTER(C.c_double)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libdgx.so not built ({LIB_PATH}); run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    L.dgx_version.restype = C.c_char_p
    sig = {
        "dgx_face_permutation": [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)],
        "dgx_reference_nodes": [C.c_int, C.c_int, dp, dp, C.POINTER(C.c_int)],
        "dgx_tri_rule": [C.c_int, C.c_int, dp, dp, C.POINTER(C.c_int)],
        "dgx_upload_mesh_mixed": [vp, C.c_void_p],
        "dgx_device_count": [C.POINTER(C.c_int)],
        "dgx_create": [C.POINTER(vp), C.POINTER(DgxConfig)],
        "dgx_destroy": [vp],
        "dgx_set_stream": [vp, vp],
        "dgx_synchronize": [vp],
        "dgx_sizes": [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_longlong)],
        "dgx_upload_mesh": [vp, C.POINTER(DgxMesh)],
        "dgx_upload_mesh_hp": [vp, C.POINTER(DgxMeshHp)],
        "dgx_get_geometry": [vp, dp, dp, dp],
        "dgx_compute_rhs": [vp, vp, vp],
        "dgx_compute_volume_rhs": [vp, vp, vp],
        "dgx_compute_face_rhs": [vp, vp, vp],
        "dgx_compute_rhs_f32": [vp, vp, vp],
        "dgx_compute_rhs_mixed": [vp, vp, vp],
        "dgx_compute_rhs_explicit": [vp, vp, vp, dp],
        "dgx_apply_minv": [vp, vp],
        "dgx_enforce_positivity": [vp, vp],
        "dgx_ssp_rk3_step": [vp, vp, C.c_double, C.c_int, dp],
        "dgx_ssp_rk3_step_noalloc": [vp, vp, C.c_double, C.c_int, dp],
        "dgx_graph_create_rhs": [vp, C.c_int, vp, vp, C.POINTER(vp)],
        "dgx_graph_create_rk3": [vp, vp, C.c_double, C.c_int, C.POINTER(vp)],
        "dgx_graph_launch": [vp, C.c_int],
        "dgx_graph_sync": [vp, dp],
        "dgx_graph_destroy": [vp],
        "dgx_set_rhs_grid": [vp, C.c_int, C.c_int],
        "dgx_autotune_rhs": [vp, C.c_int, vp, vp, C.POINTER(C.c_int), dp],
        "dgx_max_abs": [vp, vp, C.c_longlong, dp],
        "dgx_convert_f64_f32": [vp, vp, vp, C.c_longlong],
        "dgx_convert_f32_f64": [vp, vp, vp, C.c_longlong],
        "dgx_assemble_diag_blocks": [vp, vp, vp, vp],
        "dgx_cfl_dt": [vp, vp, C.c_double, vp],
        "dgx_mass_sub": [vp, vp, vp, vp],
        "dgx_solve_implicit": [vp, vp, C.POINTER(ImplicitParams), C.POINTER(ImplicitStats)],
        "dgx_jacobian_vector": [vp, vp, vp, vp],
        "dgx_assemble_face_blocks": [vp, vp, vp],
        "dgx_jacobian_bsr_pattern": [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_longlong)],
        "dgx_assemble_jacobian_bsr": [vp, vp, vp, vp],
        "dgx_newton_operator": [vp, vp, vp, vp, vp],
        "dgx_precond_apply": [vp, C.c_int, vp, vp, vp, vp],
        "dgx_mesh_read_gmsh": [C.c_char_p, C.c_int, C.c_void_p],
        "dgx_mesh_read_gmsh_mixed": [C.c_char_p, C.c_int, C.c_void_p],
        "dgx_restart_write_gmsh": [C.c_char_p, C.POINTER(DgxConfig), C.POINTER(DgxMesh), dp],
        "dgx_restart_read_solution": [C.c_char_p, C.POINTER(DgxConfig), C.c_int, dp],
        "dgx_mlp_create": [C.POINTER(vp), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_ulonglong],
        "dgx_mlp_destroy": [vp],
        "dgx_mlp_sizes": [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_longlong)],
        "dgx_mlp_get_params": [vp, C.POINTER(C.c_float)],
        "dgx_mlp_set_params": [vp, C.POINTER(C.c_float)],
        "dgx_mlp_features": [vp, vp, vp, vp, vp, vp],
        "dgx_mlp_loss_grad": [vp, vp, vp, vp, C.c_int, dp, C.POINTER(C.c_float)],
        "dgx_mlp_train": [vp, vp, vp, vp, C.c_longlong, C.c_int, C.c_int, C.c_float, dp],
        "dgx_mlp_predict": [vp, vp, C.c_int, vp],
        "dgx_attach_mlp": [vp, vp],
        "dgx_block_lu_factor": [vp, vp, vp, C.c_int, C.c_int],
        "dgx_block_lu_solve": [vp, vp, vp, vp, C.c_int, C.c_int],
        "dgx_block_inverse": [vp, vp, C.c_int, C.c_int],
        "dgx_block_inv_apply": [vp, vp, vp, vp, C.c_int, C.c_int],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    _lib = L
    return L


def check(fn: str, code: int):
    if code != DGX_OK:
        raise DgxError(fn, code)


class Context:
    """One DG discretisation on the device (owns geometry; state buffers are the caller's)."""

    def __init__(self, d: int, p: int, pde: int, num_flux: int, b_adv=(0.0, 0.0, 0.0), bc_prim=None,
                 mu: float = 0.0, Pr: float = 0.72, walls=None, sutherland: bool = False, r_s: float = 1.0):
        """pde PDE_NAVIER_STOKES: viscosity mu (constant, or with sutherland the reference value of Sutherland's
        formula at T = p / (rho r_s)) and Prandtl number Pr (2-D); walls = (nf_E, omega, rho_b, E_b), the no-slip
        wall data (dgx_config wall_nf_E, wall_omega, wall_rho, wall_E)."""
        L = load()
        cfg = DgxConfig()
        cfg.d, cfg.p, cfg.pde, cfg.num_flux = d, p, pde, num_flux
        cfg.mu, cfg.Pr = float(mu), float(Pr)
        cfg.viscosity_type, cfg.r_s = (1 if sutherland else 0), float(r_s)
        if walls is not None:
            cfg.wall_nf_E, cfg.wall_omega, cfg.wall_rho, cfg.wall_E = (float(w) for w in walls)
        for k, v in enumerate(b_adv):
            cfg.b_adv[k] = float(v)
        if bc_prim is not None:
            for k, v in enumerate(bc_prim):
                cfg.bc_prim[k] = float(v)
        h = vp()
        check("dgx_create", L.dgx_create(C.byref(h), C.byref(cfg)))
        self.h = h
        self.L = L
        self.d, self.p = d, p
        Np, nv, Nfc, nd = C.c_int(), C.c_int(), C.c_int(), C.c_longlong()
        check("dgx_sizes", L.dgx_sizes(h, C.byref(Np), C.byref(nv), C.byref(Nfc), C.byref(nd)))
        self.Np, self.nv, self.Nfc = Np.value, nv.value, Nfc.value
        self.E = self.F = 0

    def close(self):
        if getattr(self, "h", None):
            self.L.dgx_destroy(self.h)
            self.h = None

    __del__ = close

    @property
    def ndof(self) -> int:
        nd = C.c_longlong(0)
        check("dgx_sizes", self.L.dgx_sizes(self.h, None, None, None, C.byref(nd)))
        return int(nd.value) if nd.value > 0 else self.E * self.Np * self.nv

    def set_stream(self, stream_ptr: int):
        check("dgx_set_stream", self.L.dgx_set_stream(self.h, vp(stream_ptr)))

    def synchronize(self):
        check("dgx_synchronize", self.L.dgx_synchronize(self.h))

    def upload_mesh(self, xyz: np.ndarray, faces: np.ndarray):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        faces = np.ascontiguousarray(faces, dtype=np.int32)
        m = DgxMesh()
        m.E, m.F = xyz.shape[0], faces.shape[0]
        m.xyz = xyz.ctypes.data_as(dp)
        m.faces = faces.ctypes.data_as(C.POINTER(C.c_int))
        check("dgx_upload_mesh", self.L.dgx_upload_mesh(self.h, C.byref(m)))
        self.E, self.F = m.E, m.F

    def upload_mesh_hp(self, mesh):
        """Variable-order 2-D mesh (mesh.HpMesh: orders non-decreasing, per-element geometry nodes)."""
        p = np.ascontiguousarray(mesh.p, dtype=np.int32)
        xyz = mesh.flat_xyz()
        faces = np.ascontiguousarray(mesh.faces, dtype=np.int32)
        m = DgxMeshHp()
        m.E, m.F = int(p.size), int(faces.shape[0])
        m.p = p.ctypes.data_as(C.POINTER(C.c_int))
        m.xyz = xyz.ctypes.data_as(dp)
        m.faces = faces.ctypes.data_as(C.POINTER(C.c_int))
        check("dgx_upload_mesh_hp", self.L.dgx_upload_mesh_hp(self.h, C.byref(m)))
        self.E, self.F = m.E, m.F

    def upload_mesh_mixed(self, mesh):
        """Mixed TRI / QUAD 2-D mesh (mesh.MixedMesh: sorted by (type, p), per-element geometry nodes)."""
        kind = np.ascontiguousarray(mesh.kind, dtype=np.int32)
        p = np.ascontiguousarray(mesh.p, dtype=np.int32)
        xyz = mesh.flat_xyz()
        faces = np.ascontiguousarray(mesh.faces, dtype=np.int32)
        m = DgxMeshMixed()
        m.E, m.F = int(p.size), int(faces.shape[0])
        m.type = kind.ctypes.data_as(C.POINTER(C.c_int))
        m.p = p.ctypes.data_as(C.POINTER(C.c_int))
        m.xyz = xyz.ctypes.data_as(dp)
        m.faces = faces.ctypes.data_as(C.POINTER(C.c_int))
        check("dgx_upload_mesh_mixed", self.L.dgx_upload_mesh_mixed(self.h, C.byref(m)))
        self.E, self.F = m.E, m.F

    def geometry(self):
        met = np.zeros(self.E * self.d * self.d * self.Np)
        dj = np.zeros(self.E * self.Np)
        nJ = np.zeros(max(self.F, 1) * self.Nfc * self.d)
        check("dgx_get_geometry", self.L.dgx_get_geometry(self.h, met.ctypes.data_as(dp), dj.ctypes.data_as(dp),
                                                          nJ.ctypes.data_as(dp)))
        return met, dj, nJ[: self.F * self.Nfc * self.d]

    # --- device-pointer entry points
    def compute_rhs(self, u: int, r: int):
        check("dgx_compute_rhs", self.L.dgx_compute_rhs(self.h, vp(u), vp(r)))

    def compute_volume_rhs(self, u: int, r: int):
        """r += the volume term (compute_volume_rlhs_dg)."""
        check("dgx_compute_volume_rhs", self.L.dgx_compute_volume_rhs(self.h, vp(u), vp(r)))

    def compute_face_rhs(self, u: int, r: int):
        """r += the face term (compute_face_rlhs_dg)."""
        check("dgx_compute_face_rhs", self.L.dgx_compute_face_rhs(self.h, vp(u), vp(r)))

    def compute_rhs_f32(self, u: int, r: int):
        check("dgx_compute_rhs_f32", self.L.dgx_compute_rhs_f32(self.h, vp(u), vp(r)))

    def compute_rhs_mixed(self, u: int, r: int):
        check("dgx_compute_rhs_mixed", self.L.dgx_compute_rhs_mixed(self.h, vp(u), vp(r)))

    def compute_rhs_explicit(self, u: int, r: int) -> float:
        mx = C.c_double()
        check("dgx_compute_rhs_explicit", self.L.dgx_compute_rhs_explicit(self.h, vp(u), vp(r), C.byref(mx)))
        return mx.value

    def apply_minv(self, r: int):
        check("dgx_apply_minv", self.L.dgx_apply_minv(self.h, vp(r)))

    def enforce_positivity(self, u: int) -> int:
        return self.L.dgx_enforce_positivity(self.h, vp(u))

    def ssp_rk3_step(self, u: int, dt: float, positivity: bool = True) -> float:
        mx = C.c_double()
        check("dgx_ssp_rk3_step", self.L.dgx_ssp_rk3_step(self.h, vp(u), dt, int(positivity), C.byref(mx)))
        return mx.value

    def graph_rhs(self, kind: int, u: int, r: int) -> "Graph":
        """One residual evaluation recorded as a HIP graph (kind GRAPH_RHS_FP64 / _FP32 / _MIXED)."""
        g = vp()
        check("dgx_graph_create_rhs", self.L.dgx_graph_create_rhs(self.h, kind, vp(u), vp(r), C.byref(g)))
        return Graph(self, g)

    def graph_rk3(self, u: int, dt: float, positivity: bool = True) -> "Graph":
        """One SSP-RK3 step (as ssp_rk3_step) recorded as a HIP graph."""
        g = vp()
        check("dgx_graph_create_rk3", self.L.dgx_graph_create_rk3(self.h, vp(u), dt, int(positivity), C.byref(g)))
        return Graph(self, g)

    def set_rhs_grid(self, kind: int, grid: int):
        check("dgx_set_rhs_grid", self.L.dgx_set_rhs_grid(self.h, kind, grid))

    def autotune_rhs(self, kind: int, u: int, r: int) -> tuple[int, float]:
        """Picks the residual launch grid of `kind` (GRAPH_RHS_* or TUNE_RHS_MINV); returns (grid, ms)."""
        g, ms = C.c_int(), C.c_double()
        check("dgx_autotune_rhs", self.L.dgx_autotune_rhs(self.h, kind, vp(u), vp(r), C.byref(g), C.byref(ms)))
        return g.value, ms.value

    def max_abs(self, x: int, n: int) -> float:
        out = C.c_double()
        check("dgx_max_abs", self.L.dgx_max_abs(self.h, vp(x), n, C.byref(out)))
        return out.value

    def convert_f64_f32(self, src: int, dst: int, n: int):
        check("dgx_convert_f64_f32", self.L.dgx_convert_f64_f32(self.h, vp(src), vp(dst), n))

    def convert_f32_f64(self, src: int, dst: int, n: int):
        check("dgx_convert_f32_f64", self.L.dgx_convert_f32_f64(self.h, vp(src), vp(dst), n))

    def attach_mlp(self, mlp: "MLP | None"):
        check("dgx_attach_mlp", self.L.dgx_attach_mlp(self.h, mlp.h if mlp is not None else None))

    def assemble_diag_blocks(self, u: int, mass_scale: int | None, blocks: int):
        """mass_scale: device pointer to E doubles (1/dt per element) or None."""
        check("dgx_assemble_diag_blocks",
              self.L.dgx_assemble_diag_blocks(self.h, vp(u), vp(mass_scale or 0), vp(blocks)))

    def newton_operator(self, u: int, mass_scale: int, v: int, out: int):
        check("dgx_newton_operator", self.L.dgx_newton_operator(self.h, vp(u), vp(mass_scale), vp(v), vp(out)))

    def precond_apply(self, precond: int, u: int, inv_dt: int, v: int, z: int):
        check("dgx_precond_apply", self.L.dgx_precond_apply(self.h, precond, vp(u), vp(inv_dt), vp(v), vp(z)))

    def assemble_face_blocks(self, u: int, blocks: int):
        check("dgx_assemble_face_blocks", self.L.dgx_assemble_face_blocks(self.h, vp(u), vp(blocks)))

    def jacobian_bsr_pattern(self):
        """(row_ptr, col_idx) of the block-sparse Newton matrix (numpy int32 arrays)."""
        import numpy as np

        n = C.c_longlong(0)
        check("dgx_jacobian_bsr_pattern", self.L.dgx_jacobian_bsr_pattern(self.h, None, None, C.byref(n)))
        rp = np.zeros(self.E + 1, dtype=np.int32)
        ci = np.zeros(max(n.value, 1), dtype=np.int32)
        check("dgx_jacobian_bsr_pattern",
              self.L.dgx_jacobian_bsr_pattern(self.h, rp.ctypes.data_as(C.POINTER(C.c_int)),
                                              ci.ctypes.data_as(C.POINTER(C.c_int)), C.byref(n)))
        return rp, ci[: n.value]

    def assemble_jacobian_bsr(self, u: int, mass_scale: int | None, values: int):
        check("dgx_assemble_jacobian_bsr",
              self.L.dgx_assemble_jacobian_bsr(self.h, vp(u), vp(mass_scale or 0), vp(values)))

    def cfl_dt(self, u: int, cfl: float, inv_dt: int):
        check("dgx_cfl_dt", self.L.dgx_cfl_dt(self.h, vp(u), cfl, vp(inv_dt)))

    def mass_sub(self, x: int, scale: int | None, y: int):
        check("dgx_mass_sub", self.L.dgx_mass_sub(self.h, vp(x), vp(scale or 0), vp(y)))

    def solve_implicit(self, u: int, cfl_initial=10.0, exit_tol=1e-10, exit_ratio=0.0, max_newton=100,
                       gmres_restart=60, gmres_rtol=1e-4, gmres_max_iter=600, precond=PC_BLOCK_JACOBI,
                       raise_on_error=True, capture=None, cfl_max=0.0, learned_fallback_iters=0, jvp_fp32=False):
        """capture: optional (blocks_ptr, features_ptr, sigma_ptr, capacity_elements) device buffers; jvp_fp32: the
        GMRES operator in FP32 dual numbers (the FP64 residual still decides convergence)."""
        cb, cf, cs, cc = capture if capture else (0, 0, 0, 0)
        p = ImplicitParams(cfl_initial, exit_tol, exit_ratio, max_newton, gmres_restart, gmres_rtol, gmres_max_iter,
                           precond, cfl_max, learned_fallback_iters, cb or None, cf or None, cs or None, cc,
                           1 if jvp_fp32 else 0)
        st = ImplicitStats()
        rc = self.L.dgx_solve_implicit(self.h, vp(u), C.byref(p), C.byref(st))
        if raise_on_error and rc != DGX_OK:
            raise DgxError("dgx_solve_implicit", rc)
        out = {k: getattr(st, k) for k in ("newton_steps", "gmres_iters", "gmres_not_converged", "converged",
                                           "max_rhs0", "max_rhs", "seconds", "captured", "fallback_steps")}
        out["history"] = list(st.history[: min(st.newton_steps + 1, HISTORY_MAX)])
        out["gmres_per_step"] = list(st.gmres_per_step[: min(st.newton_steps, HISTORY_MAX)])
        out["status"] = rc
        return out

    def jacobian_vector(self, u: int, v: int, jv: int):
        check("dgx_jacobian_vector", self.L.dgx_jacobian_vector(self.h, vp(u), vp(v), vp(jv)))

    def block_lu_factor(self, blocks: int, piv: int, E: int, nb: int):
        check("dgx_block_lu_factor", self.L.dgx_block_lu_factor(self.h, vp(blocks), vp(piv), E, nb))

    def block_inverse(self, blocks: int, E: int, nb: int):
        check("dgx_block_inverse", self.L.dgx_block_inverse(self.h, vp(blocks), E, nb))

    def block_inv_apply(self, inv: int, x: int, y: int, E: int, nb: int):
        check("dgx_block_inv_apply", self.L.dgx_block_inv_apply(self.h, vp(inv), vp(x), vp(y), E, nb))

    def block_lu_solve(self, lu: int, piv: int, x: int, E: int, nb: int):
        check("dgx_block_lu_solve", self.L.dgx_block_lu_solve(self.h, vp(lu), vp(piv), vp(x), E, nb))


MLP_LOWRANK, MLP_KRON = 0, 1


GRAPH_RHS_FP64, GRAPH_RHS_FP32, GRAPH_RHS_MIXED, TUNE_RHS_MINV = 0, 1, 2, 3


class Graph:
    """A recorded launch sequence (dgx_graph_*); keeps its context alive."""

    def __init__(self, ctx: Context, h):
        self.ctx, self.L, self.h = ctx, ctx.L, h

    def launch(self, n: int = 1):
        check("dgx_graph_launch", self.L.dgx_graph_launch(self.h, n))

    def sync(self) -> float:
        mx = C.c_double()
        check("dgx_graph_sync", self.L.dgx_graph_sync(self.h, C.byref(mx)))
        return mx.value

    def close(self):
        if getattr(self, "h", None):
            self.L.dgx_graph_destroy(self.h)
            self.h = None

    __del__ = close


class MLP:
    """The learned block preconditioner (dgx_mlp_*)."""

    def __init__(self, nb: int, Np: int, hidden: int = 256, rank: int | None = None, seed: int = 20261202,
                 kind: int = 0):
        self.L = load()
        h = vp()
        if rank is None:
            rank = 8
        check("dgx_mlp_create", self.L.dgx_mlp_create(C.byref(h), kind, nb, Np, hidden, rank, seed))
        self.kind = kind
        self.h = h
        nf, nout, npar = C.c_int(), C.c_int(), C.c_longlong()
        check("dgx_mlp_sizes", self.L.dgx_mlp_sizes(h, C.byref(nf), C.byref(nout), C.byref(npar)))
        self.nf, self.nout, self.nparam = nf.value, nout.value, npar.value
        self.nb, self.Np, self.hidden, self.rank = nb, Np, hidden, rank

    def close(self):
        if getattr(self, "h", None):
            self.L.dgx_mlp_destroy(self.h)
            self.h = None

    __del__ = close

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.nparam, dtype=np.float32)
        check("dgx_mlp_get_params", self.L.dgx_mlp_get_params(self.h, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def set_params(self, w: np.ndarray):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == self.nparam
        check("dgx_mlp_set_params", self.L.dgx_mlp_set_params(self.h, w.ctypes.data_as(C.POINTER(C.c_float))))

    def features(self, ctx: "Context", u: int, inv_dt: int, feat: int, sigma: int):
        check("dgx_mlp_features", self.L.dgx_mlp_features(ctx.h, self.h, vp(u), vp(inv_dt), vp(feat), vp(sigma)))

    def loss_grad(self, feat: int, sigma: int | None, blocks: int, B: int, want_grad=True):
        loss = C.c_double()
        g = np.zeros(self.nparam, dtype=np.float32) if want_grad else None
        check("dgx_mlp_loss_grad", self.L.dgx_mlp_loss_grad(
            self.h, vp(feat), vp(sigma or 0), vp(blocks), B, C.byref(loss),
            g.ctypes.data_as(C.POINTER(C.c_float)) if want_grad else None))
        return loss.value, g

    def train(self, feat: int, sigma: int | None, blocks: int, n: int, epochs: int, batch: int, lr: float) -> float:
        loss = C.c_double()
        check("dgx_mlp_train", self.L.dgx_mlp_train(self.h, vp(feat), vp(sigma or 0), vp(blocks), n, epochs, batch,
                                                   lr, C.byref(loss)))
        return loss.value

    def predict(self, feat: int, B: int, out: int):
        check("dgx_mlp_predict", self.L.dgx_mlp_predict(self.h, vp(feat), B, vp(out)))

    def attach(self, ctx: "Context"):
        check("dgx_attach_mlp", self.L.dgx_attach_mlp(ctx.h, self.h))


def face_permutation(d: int, p: int, ord_: int) -> np.ndarray:
    """dgx_face_permutation (host only): perm[i] = the right-face node that coincides with left-face node i."""
    L = load()
    n = (p + 1) ** (d - 1)
    out = np.zeros(n, dtype=np.int32)
    check("dgx_face_permutation", L.dgx_face_permutation(d, p, ord_, out.ctypes.data_as(C.POINTER(C.c_int))))
    return out


def reference_nodes(kind: int, p: int):
    """dgx_reference_nodes (host only): (rst (Np, 2), w (Np,)) of the solution nodes of type kind (0 QUAD, 1 TRI)."""
    L = load()
    n = C.c_int()
    check("dgx_reference_nodes", L.dgx_reference_nodes(kind, p, None, None, C.byref(n)))
    rst, w = np.zeros(2 * n.value), np.zeros(n.value)
    check("dgx_reference_nodes", L.dgx_reference_nodes(kind, p, rst.ctypes.data_as(dp), w.ctypes.data_as(dp), None))
    return rst.reshape(-1, 2), w


def tri_rule(kind: int, q: int):
    """dgx_tri_rule (host only): the reference's TRI rule (kind 0 WSH order q, 1 WV degree q): (rst (n, 2), w)."""
    L = load()
    n = C.c_int()
    check("dgx_tri_rule", L.dgx_tri_rule(kind, q, None, None, C.byref(n)))
    rst, w = np.zeros(2 * n.value), np.zeros(n.value)
    check("dgx_tri_rule", L.dgx_tri_rule(kind, q, rst.ctypes.data_as(dp), w.ctypes.data_as(dp), None))
    return rst.reshape(-1, 2), w


def read_gmsh(path: str, p: int) -> dict:
    """dgx_mesh_read_gmsh: a gmsh 2.2 QUAD mesh as numpy arrays (xyz, faces, v_to_v, v_to_lf, v_to_lf_wp)."""
    L = load()
    L.dgx_mesh_free.argtypes = [C.c_void_p]
    L.dgx_mesh_free.restype = None
    m = DgxHostMesh()
    check("dgx_mesh_read_gmsh", L.dgx_mesh_read_gmsh(str(path).encode(), p, C.byref(m)))
    try:
        Np = (p + 1) ** m.d
        nf = 2 * m.d
        out = {"d": m.d, "p": m.p, "E": m.E, "F": m.F,
               "xyz": np.ctypeslib.as_array(m.xyz, (m.E, m.d, Np)).copy(),
               "faces": np.ctypeslib.as_array(m.faces, (m.F, 6)).copy(),
               "v_to_v": np.ctypeslib.as_array(m.v_to_v, (m.E, nf)).copy(),
               "v_to_lf": np.ctypeslib.as_array(m.v_to_lf, (m.E, nf)).copy(),
               "v_to_lf_wp": np.ctypeslib.as_array(m.v_to_lf_wp, (m.E, nf)).copy(),
               "n_unsupported_bc": m.n_unsupported_bc, "n_inverted": m.n_inverted}
    finally:
        L.dgx_mesh_free(C.byref(m))
    return out


def read_gmsh_mixed(path: str, p: int) -> dict:
    """dgx_mesh_read_gmsh_mixed: a gmsh 2.2 TRI / QUAD / mixed mesh, volumes in file order (type, xyz as a list
    of (2, Np_e) arrays, faces, and v_to_v / v_to_lf / v_to_lf_wp with 4 entries per volume, a TRI's 4th -1)."""
    L = load()
    L.dgx_mesh_mixed_free.argtypes = [C.c_void_p]
    L.dgx_mesh_mixed_free.restype = None
    L.dgx_mesh_read_gmsh_mixed.argtypes = [C.c_char_p, C.c_int, C.c_void_p]
    L.dgx_mesh_read_gmsh_mixed.restype = C.c_int
    m = DgxHostMeshMixed()
    check("dgx_mesh_read_gmsh_mixed", L.dgx_mesh_read_gmsh_mixed(str(path).encode(), p, C.byref(m)))
    try:
        typ = np.ctypeslib.as_array(m.type, (m.E,)).copy()
        npe = np.where(typ == 1, (p + 1) * (p + 2) // 2, (p + 1) ** 2)
        flat = np.ctypeslib.as_array(m.xyz, (int(2 * npe.sum()),)).copy()
        off = np.concatenate([[0], np.cumsum(2 * npe)])
        out = {"p": m.p, "E": m.E, "F": m.F, "type": typ,
               "xyz": [flat[off[e]:off[e + 1]].reshape(2, -1) for e in range(m.E)],
               "faces": np.ctypeslib.as_array(m.faces, (m.F, 6)).copy(),
               "v_to_v": np.ctypeslib.as_array(m.v_to_v, (m.E, 4)).copy(),
               "v_to_lf": np.ctypeslib.as_array(m.v_to_lf, (m.E, 4)).copy(),
               "v_to_lf_wp": np.ctypeslib.as_array(m.v_to_lf_wp, (m.E, 4)).copy(),
               "n_unsupported_bc": m.n_unsupported_bc, "n_inverted": m.n_inverted}
    finally:
        L.dgx_mesh_mixed_free(C.byref(m))
    return out


def _config(d, p, pde, num_flux=0):
    cfg = DgxConfig()
    cfg.d, cfg.p, cfg.pde, cfg.num_flux = d, p, pde, num_flux
    return cfg


def restart_write(path: str, d: int, p: int, pde: int, xyz: np.ndarray, faces: np.ndarray, u: np.ndarray):
    """dgx_restart_write_gmsh for a host mesh and host state."""
    L = load()
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    faces = np.ascontiguousarray(faces, dtype=np.int32)
    u = np.ascontiguousarray(u, dtype=np.float64)
    m = DgxMesh(xyz.shape[0], faces.shape[0], xyz.ctypes.data_as(C.POINTER(C.c_double)),
                faces.ctypes.data_as(C.POINTER(C.c_int)))
    cfg = _config(d, p, pde)
    check("dgx_restart_write_gmsh", L.dgx_restart_write_gmsh(str(path).encode(), C.byref(cfg), C.byref(m),
                                                             u.ctypes.data_as(C.POINTER(C.c_double))))


def restart_read(path: str, d: int, p: int, pde: int, E: int) -> np.ndarray:
    """dgx_restart_read_solution: nodal state from a restart file."""
    L = load()
    nv = d + 2 if pde == 1 else 1
    u = np.zeros(E * nv * (p + 1) ** d)
    cfg = _config(d, p, pde)
    check("dgx_restart_read_solution", L.dgx_restart_read_solution(str(path).encode(), C.byref(cfg), E,
                                                                   u.ctypes.data_as(C.POINTER(C.c_double))))
    return u
