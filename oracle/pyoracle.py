"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.or_create.restype = C.c_void_p
        L.or_create.argtypes = [C.c_int] * 6 + [dp, ip, dp, dp]
        L.or_destroy.argtypes = [C.c_void_p]
        for name in ("or_compute_rhs_d", "or_volume_rhs_d", "or_face_rhs_d"):
            getattr(L, name).argtypes = [C.c_void_p, dp, dp]
        L.or_compute_rhs_d_threads.argtypes = [C.c_void_p, dp, dp, C.c_int]
        L.or_set_viscous.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.or_set_sutherland.argtypes = [C.c_void_p, C.c_double]
        L.or_set_walls.argtypes = [C.c_void_p] + [C.c_double] * 4
        L.or_viscous_flux.argtypes = [C.c_int, dp, dp, C.c_double, C.c_double, dp]
        L.or_ns_gradient.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.or_ns_gradient.restype = C.c_int
        L.or_compute_rhs_ns.argtypes = [C.c_void_p, dp, dp]
        L.or_compute_rhs_ns.restype = C.c_int
        L.or_compute_rhs_ns_c.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.or_compute_rhs_ns_c.restype = C.c_int
        L.or_compute_rhs_c.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.or_geometry.argtypes = [C.c_void_p, dp, dp, dp]
        L.or_apply_minv.argtypes = [C.c_void_p, dp]
        L.or_mass.argtypes = [C.c_void_p, C.c_int, dp]
        L.or_compute_rhs_explicit.argtypes = [C.c_void_p, dp, dp]
        L.or_compute_rhs_explicit.restype = C.c_double
        L.or_enforce_positivity.argtypes = [C.c_void_p, dp]
        L.or_ssp_rk3_step.argtypes = [C.c_void_p, dp, C.c_double, C.c_int, ip]
        L.or_ssp_rk3_step.restype = C.c_double
        L.or_gauss_legendre.argtypes = [C.c_int, dp, dp]
        L.or_gauss_lobatto.argtypes = [C.c_int, dp, dp]
        L.or_lagrange.argtypes = [C.c_int, dp, C.c_int, dp, dp, dp]
        L.or_kron.argtypes = [C.c_int, C.POINTER(dp), ip, ip, dp]
        L.or_bernstein.argtypes = [C.c_int, C.c_int, dp, dp]
        L.or_euler_flux_d.argtypes = [C.c_int, dp, dp]
        L.or_numflux_d.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp, dp, dp, dp]
        L.or_numflux_c.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp, dp, dp, dp, dp, dp, dp]
        L.or_bc_state_d.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp]
        L.or_jacobian_block_cs.argtypes = [C.c_void_p, dp, C.c_int, C.c_int, C.c_double, dp]
        L.or_diag_blocks_cs.argtypes = [C.c_void_p, dp, C.c_double, dp]
        L.or_diag_blocks_cs_threads.argtypes = [C.c_void_p, dp, C.c_double, dp, C.c_int]
        L.or_compute_rhs_c_threads.argtypes = [C.c_void_p, dp, dp, dp, dp, C.c_int]
        L.or_cfl_inv_dt.argtypes = [C.c_void_p, dp, C.c_double, dp]
        L.or_relaxed_update.argtypes = [C.c_void_p, dp, dp]
        L.or_relaxed_update.restype = C.c_int
        L.or_block_apply.argtypes = [C.c_int, C.c_int, dp, dp, dp, C.c_int]
        L.or_block_inverse.argtypes = [C.c_int, C.c_int, dp, dp, C.c_int]
        L.or_block_inverse.restype = C.c_int
        L.or_dots.argtypes = [C.c_int, C.c_longlong, dp, dp, dp, C.c_int]
        L.or_gemv_sub.argtypes = [C.c_int, C.c_longlong, dp, dp, dp, C.c_int]
        L.or_face_perm.argtypes = [C.c_void_p, C.c_int, ip]
        _lib = L
    return _lib


def _d(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(dp)


def _i(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(ip)


def gauss_legendre(n):
    x, w = np.zeros(n), np.zeros(n)
    lib().or_gauss_legendre(n, _d(x), _d(w))
    return x, w


def gauss_lobatto(n):
    x, w = np.zeros(n), np.zeros(n)
    lib().or_gauss_lobatto(n, _d(x), _d(w))
    return x, w


def lagrange(xin, xout):
    xin = np.ascontiguousarray(xin, dtype=np.float64)
    xout = np.ascontiguousarray(xout, dtype=np.float64)
    I = np.zeros((len(xout), len(xin)))
    D = np.zeros((len(xout), len(xin)))
    lib().or_lagrange(len(xin), _d(xin), len(xout), _d(xout), _d(I), _d(D))
    return I, D


def kron(ops):
    ops = [np.ascontiguousarray(o, dtype=np.float64) for o in ops]
    rows = np.array([o.shape[0] for o in ops], dtype=np.int32)
    cols = np.array([o.shape[1] for o in ops], dtype=np.int32)
    out = np.zeros((int(np.prod(rows)), int(np.prod(cols))))
    arr = (dp * len(ops))(*[_d(o) for o in ops])
    lib().or_kron(len(ops), arr, _i(rows), _i(cols), _d(out))
    return out


def euler_flux(d, s):
    s = np.ascontiguousarray(s, dtype=np.float64)
    f = np.zeros(d * (d + 2))
    lib().or_euler_flux_d(d, _d(s), _d(f))
    return f


def numflux(d, pde, flux, sL, sR, n, b=(0.0, 0.0, 0.0)):
    nv = d + 2 if pde == 1 else 1
    out = np.zeros(nv)
    lib().or_numflux_d(d, pde, flux, _d(np.ascontiguousarray(sL, float)), _d(np.ascontiguousarray(sR, float)),
                       _d(np.ascontiguousarray(n, float)), _d(np.ascontiguousarray(b, float)), _d(out))
    return out


def numflux_cs(d, pde, flux, sL, sR, n, dsL, dsR, h=1e-30, b=(0.0, 0.0, 0.0)):
    """Im(F(sL + i h dsL, sR + i h dsR)) / h."""
    nv = d + 2 if pde == 1 else 1
    re, im = np.zeros(nv), np.zeros(nv)
    f = lambda a: _d(np.ascontiguousarray(a, dtype=np.float64))
    lib().or_numflux_c(d, pde, flux, f(sL), f(h * np.asarray(dsL)), f(sR), f(h * np.asarray(dsR)), f(n), f(b),
                       _d(re), _d(im))
    return re, im / h


def bc_state(d, bc, sL, n, prim):
    out = np.zeros(d + 2)
    f = lambda a: _d(np.ascontiguousarray(a, dtype=np.float64))
    lib().or_bc_state_d(d, bc, f(sL), f(n), f(np.resize(np.asarray(prim, float), 5)), _d(out))
    return out


class Oracle:
    """One DG discretisation (mesh + physics) on the CPU."""

    def __init__(self, mesh, pde: int, flux: int, b_adv=(0.0, 0.0, 0.0), bc_prim=None, viscous=None, walls=None,
                 sutherland_r_s=None):
        """viscous = (mu, Pr): Navier-Stokes with BR2 (Euler contexts; interior, Riemann and no-slip wall faces).
        walls = (nf_E, omega, rho_b, E_b): the no-slip wall data (or_set_walls)."""
        self.mesh = mesh
        self.d, self.p = mesh.d, mesh.p
        self.pde, self.flux = pde, flux
        self.nv = mesh.d + 2 if pde == 1 else 1
        self.Np = (mesh.p + 1) ** mesh.d
        self.Nfc = (mesh.p + 1) ** (mesh.d - 1)
        xyz = np.ascontiguousarray(mesh.xyz, dtype=np.float64)
        faces = np.ascontiguousarray(mesh.faces, dtype=np.int32)
        b = np.zeros(3)
        b[: len(b_adv)] = b_adv
        bp = np.zeros(5)
        if bc_prim is not None:
            bp[: len(bc_prim)] = bc_prim
        self._keep = (xyz, faces, b, bp)
        self.ctx = lib().or_create(mesh.d, mesh.p, pde, flux, mesh.E, mesh.F, _d(xyz), _i(faces), _d(b), _d(bp))
        if not self.ctx:
            raise RuntimeError("or_create failed")
        self.ndof = mesh.E * self.Np * self.nv
        if viscous is not None:
            lib().or_set_viscous(self.ctx, float(viscous[0]), float(viscous[1]))
            if sutherland_r_s is not None:
                lib().or_set_sutherland(self.ctx, float(sutherland_r_s))
        if walls is not None:
            lib().or_set_walls(self.ctx, *[float(w) for w in walls])

    def __del__(self):
        if getattr(self, "ctx", None):
            lib().or_destroy(self.ctx)
            self.ctx = None

    def geometry(self):
        m = self.mesh
        met = np.zeros(m.E * m.d * m.d * self.Np)
        dj = np.zeros(m.E * self.Np)
        nJ = np.zeros(m.F * self.Nfc * m.d)
        lib().or_geometry(self.ctx, _d(met), _d(dj), _d(nJ))
        return met, dj, nJ

    def rhs(self, u, part="all"):
        u = np.ascontiguousarray(u, dtype=np.float64)
        r = np.zeros(self.ndof)
        fn = {"all": lib().or_compute_rhs_d, "volume": lib().or_volume_rhs_d, "face": lib().or_face_rhs_d}[part]
        fn(self.ctx, _d(u), _d(r))
        return r

    def rhs_threads(self, u, nthreads):
        """The residual on nthreads OpenMP threads (faces lifted by colour); equal to rhs() to rounding."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        r = np.zeros(self.ndof)
        lib().or_compute_rhs_d_threads(self.ctx, _d(u), _d(r), int(nthreads))
        return r

    def rhs_complex(self, u_re, u_im):
        rre, rim = np.zeros(self.ndof), np.zeros(self.ndof)
        lib().or_compute_rhs_c(self.ctx, _d(np.ascontiguousarray(u_re, float)),
                               _d(np.ascontiguousarray(u_im, float)), _d(rre), _d(rim))
        return rre, rim

    def apply_minv(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64).copy()
        if lib().or_apply_minv(self.ctx, _d(r)):
            raise RuntimeError("singular mass matrix")
        return r

    def rhs_ns(self, u):
        """Navier-Stokes (BR2) residual, no M^-1."""
        r = np.zeros(self.ndof)
        if lib().or_compute_rhs_ns(self.ctx, _d(np.ascontiguousarray(u, float)), _d(r)):
            raise RuntimeError("or_compute_rhs_ns: viscous context with interior, Riemann and no-slip wall faces only")
        return r

    def rhs_ns_complex(self, u_re, u_im):
        """The Navier-Stokes residual at a complex state (its imaginary part carries complex-step derivatives)."""
        rre, rim = np.zeros(self.ndof), np.zeros(self.ndof)
        if lib().or_compute_rhs_ns_c(self.ctx, _d(np.ascontiguousarray(u_re, float)),
                                     _d(np.ascontiguousarray(u_im, float)), _d(rre), _d(rim)):
            raise RuntimeError("or_compute_rhs_ns_c: viscous context with interior, Riemann and no-slip wall faces only")
        return rre, rim

    def jvp_ns(self, u, v, h=1e-30):
        """J v of the Navier-Stokes residual by complex step: Im R(u + i h v) / h."""
        return self.rhs_ns_complex(u, h * np.asarray(v, float))[1] / h

    def ns_diag_block(self, u, e, h=1e-30):
        """The diagonal block dR_e/du_e (nb x nb, nb = nv Np, rows and columns in the e x nv x Np order) by
        complex step, one local dof at a time."""
        nb = self.nv * self.Np
        B = np.empty((nb, nb))
        im = np.zeros(self.ndof)
        for col in range(nb):
            im[:] = 0.0
            im[e * nb + col] = h
            B[:, col] = self.rhs_ns_complex(u, im)[1][e * nb:(e + 1) * nb] / h
        return B

    def ns_face_block(self, u, e, lf, h=1e-30):
        """The face block dR_e/du_n (nb x nb, rows and columns in the e x nv x Np order), n the neighbour across
        local face lf of element e, by complex step, one of the neighbour's dofs at a time; zero on a boundary
        face."""
        nb = self.nv * self.Np
        f = self.mesh.faces
        hit = [(int(r[2]) if (r[0] == e and r[1] == lf) else int(r[0])) for r in f
               if (r[0] == e and r[1] == lf) or (r[2] == e and r[3] == lf)]
        B = np.zeros((nb, nb))
        if not hit or hit[0] < 0:
            return B
        n = hit[0]
        im = np.zeros(self.ndof)
        for col in range(nb):
            im[:] = 0.0
            im[n * nb + col] = h
            B[:, col] = self.rhs_ns_complex(u, im)[1][e * nb:(e + 1) * nb] / h
        return B

    def ns_gradient(self, u):
        """(g, gv, gf): full and volume-only weak gradient coefficients E x d x nv x Np, face lifts
        F x 2 x d x nv x Np."""
        m = self.mesh
        eb = m.d * self.nv * self.Np
        g, gv, gf = np.zeros(m.E * eb), np.zeros(m.E * eb), np.zeros(m.F * 2 * eb)
        if lib().or_ns_gradient(self.ctx, _d(np.ascontiguousarray(u, float)), _d(g), _d(gv), _d(gf)):
            raise RuntimeError("or_ns_gradient: interior, Riemann and no-slip wall faces only")
        sh = (m.E, m.d, self.nv, self.Np)
        return g.reshape(sh), gv.reshape(sh), gf.reshape(m.F, 2, m.d, self.nv, self.Np)

    def mass(self, e):
        M = np.zeros((self.Np, self.Np))
        lib().or_mass(self.ctx, e, _d(M))
        return M

    def rhs_explicit(self, u):
        r = np.zeros(self.ndof)
        mx = lib().or_compute_rhs_explicit(self.ctx, _d(np.ascontiguousarray(u, float)), _d(r))
        return r, mx

    def positivity(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64).copy()
        st = lib().or_enforce_positivity(self.ctx, _d(u))
        return u, st

    def ssp_rk3_step(self, u, dt, positivity=True):
        u = np.ascontiguousarray(u, dtype=np.float64).copy()
        err = np.zeros(1, dtype=np.int32)
        mx = lib().or_ssp_rk3_step(self.ctx, _d(u), dt, int(positivity), _i(err))
        return u, mx, int(err[0])

    def jacobian_block(self, u, e_row, e_col, h=1e-30):
        nb = self.Np * self.nv
        blk = np.zeros((nb, nb))
        lib().or_jacobian_block_cs(self.ctx, _d(np.ascontiguousarray(u, float)), e_row, e_col, h, _d(blk))
        return blk

    def diag_blocks(self, u, h=1e-30, nthreads=1):
        nb = self.Np * self.nv
        out = np.zeros((self.mesh.E, nb, nb))
        if nthreads > 1:
            lib().or_diag_blocks_cs_threads(self.ctx, _d(np.ascontiguousarray(u, float)), h, _d(out), int(nthreads))
        else:
            lib().or_diag_blocks_cs(self.ctx, _d(np.ascontiguousarray(u, float)), h, _d(out))
        return out

    def rhs_complex_threads(self, u_re, u_im, nthreads):
        rre, rim = np.zeros(self.ndof), np.zeros(self.ndof)
        lib().or_compute_rhs_c_threads(self.ctx, _d(np.ascontiguousarray(u_re, float)),
                                       _d(np.ascontiguousarray(u_im, float)), _d(rre), _d(rim), int(nthreads))
        return rre, rim

    def cfl_inv_dt(self, u, cfl):
        out = np.zeros(self.mesh.E)
        lib().or_cfl_inv_dt(self.ctx, _d(np.ascontiguousarray(u, float)), float(cfl), _d(out))
        return out

    def relaxed_update(self, u, du):
        """u += relax_e du_e in place (u: contiguous float64); returns the status (0, or -1: relax < EPS)."""
        return lib().or_relaxed_update(self.ctx, _d(u), _d(np.ascontiguousarray(du, float)))


def viscous_flux(d, s, g, mu, Pr):
    """Negated viscous flux fv[dim + d*eq] (or_viscous_flux)."""
    fv = np.zeros(d * (d + 2))
    lib().or_viscous_flux(d, _d(np.ascontiguousarray(s, float)), _d(np.ascontiguousarray(g, float)), mu, Pr, _d(fv))
    return fv.reshape(d + 2, d)


def block_apply(A, v, nthreads=1):
    """z_e = A_e v_e for a stack of dense blocks A (E, nb, nb) and v (E * nb,)."""
    E, nb, _ = A.shape
    z = np.zeros(E * nb)
    lib().or_block_apply(E, nb, _d(np.ascontiguousarray(A, float)), _d(np.ascontiguousarray(v, float)), _d(z),
                         int(nthreads))
    return z


def block_inverse(A, nthreads=1):
    """Inverses of a stack of dense blocks (E, nb, nb), Gauss-Jordan with partial pivoting."""
    E, nb, _ = A.shape
    out = np.zeros_like(A)
    if lib().or_block_inverse(E, nb, _d(np.ascontiguousarray(A, float)), _d(out), int(nthreads)):
        raise np.linalg.LinAlgError("singular block")
    return out


def dots(V, k, w, nthreads=1):
    """V[:k] @ w (V: (m, n) C-contiguous)."""
    h = np.zeros(k)
    lib().or_dots(k, V.shape[1], _d(V), _d(w), _d(h), int(nthreads))
    return h


def gemv_sub(V, k, h, w, nthreads=1):
    """w -= h @ V[:k] in place."""
    lib().or_gemv_sub(k, V.shape[1], _d(V), _d(np.ascontiguousarray(h, float)), _d(w), int(nthreads))
