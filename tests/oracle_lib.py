"""ctypes binding of the CPU FP64 oracle (oracle/liboracle.so).

The oracle is the checker: only tests, the smoke check and the bench's
CPU-baseline leg load it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        L = C.CDLL(LIB)
        L.orc_mesh_create.restype = C.c_void_p
        L.orc_mesh_create.argtypes = [C.c_int, C.c_int, _D, _D, _I, _I, C.c_int,
                                      C.c_double, C.c_double]
        for name in ("orc_mesh_destroy",):
            getattr(L, name).argtypes = [C.c_void_p]
        L.orc_sys_create.restype = C.c_void_p
        L.orc_sys_create.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int]
        L.orc_sys_destroy.argtypes = [C.c_void_p]
        L.orc_solve_plane.argtypes = [C.c_void_p, _D, _D, _D, C.c_double, C.c_int]
        L.orc_field_solve.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double,
                                      C.c_int, _D, _D, _D, C.c_double, C.c_int,
                                      C.c_int, C.c_int, _D]
        L.orc_field_solve_iters.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double,
                                            C.c_int, _D, _D, _D, C.c_double, C.c_int, _I]
        L.orc_field_solve_base.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double,
                                           C.c_int, C.c_int, _D, _D, _D, C.c_double,
                                           C.c_int, _I]
        L.orc_dftr.argtypes = [_D, C.c_int, C.c_int, C.c_int]
        L.orc_gradient.argtypes = [C.c_void_p, C.c_int, C.c_double, _D, C.c_int]
        L.orc_smooth.argtypes = [C.c_void_p, C.c_int, _D]
        L.orc_helm_op.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, _D, _D]
        L.orc_helm_matrix.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, _D]
        L.orc_helm_sc.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, _D, _D, _D]
        L.orc_helm_diag.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, _D]
        L.orc_helm_global.argtypes = [C.c_void_p, _D, _D]
        L.orc_svv_coeffs_z.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                       C.c_double, _D]
        L.orc_set_svv_z.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int]
        L.orc_mesh_enumerate.restype = C.c_int
        L.orc_mesh_enumerate.argtypes = [C.c_void_p, C.c_int]
        L.orc_mesh_set_cylindrical.restype = C.c_int
        L.orc_mesh_set_cylindrical.argtypes = [C.c_void_p]
        L.orc_set_field_name.restype = None
        L.orc_set_field_name.argtypes = [C.c_int]
        L.orc_dns_set_iterative.restype = C.c_int
        L.orc_dns_set_iterative.argtypes = [C.c_void_p, C.c_int]
        L.orc_dns_set_bc.restype = None
        L.orc_dns_set_bc.argtypes = [C.c_void_p, _D, _D, _D]
        L.orc_couple.restype = None
        L.orc_couple.argtypes = [_D, _D, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_mesh_renumber.restype = C.c_int
        L.orc_mesh_renumber.argtypes = [C.c_void_p, _I, _I, C.c_int, C.c_int]
        L.orc_dns_create.restype = C.c_void_p
        L.orc_dns_create.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double,
                                     C.c_double, C.c_int, C.c_int, C.c_double,
                                     C.c_int, C.c_int]
        L.orc_dns_destroy.argtypes = [C.c_void_p]
        L.orc_dns_create_bc.restype = C.c_void_p
        L.orc_dns_create_bc.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_double,
                                        C.c_double, C.c_double, C.c_int, C.c_int,
                                        C.c_double, C.c_int, C.c_int, _D, _D, _D]
        L.orc_dns_run.argtypes = [C.c_void_p, C.c_int, _D, _D, _D, _D, _D]
        _lib = L
    return _lib


def dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def ip(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_I)


def gll(np_):
    z = np.zeros(np_); w = np.zeros(np_)
    D = np.zeros((np_, np_)); DT = np.zeros((np_, np_))
    lib().orc_zwglj(dp(z), dp(w), C.c_int(np_))
    lib().orc_dglj(dp(D), dp(DT), dp(z), C.c_int(np_))
    return z, w, D, DT


def svv_operators(np_, mn=-1, eps=0.0, nu=1.0):
    S = np.zeros((np_, np_)); ST = np.zeros((np_, np_))
    lib().orc_svv_operators(C.c_int(np_), C.c_int(mn), C.c_double(eps),
                            C.c_double(nu), dp(S), dp(ST))
    return S, ST


def svv_coeffs_z(nummodes, nmode, base, mz, eps, nu):
    S = np.zeros(nummodes)
    lib().orc_svv_coeffs_z(nummodes, nmode, base, mz, eps, nu, dp(S))
    return S


def set_svv_z(mz=-1, eps=0.0, nu=1.0, nmode=0):
    """Fourier-direction SVV settings of the oracle's field solves (tokens
    SVV_MZ, SVV_EPSZ, KINVIS; nmode = Geometry::nMode, 0: base + nmodes)."""
    lib().orc_set_svv_z(mz, eps, nu, nmode)


def edgemaps(np_):
    e = np.zeros(np_ * np_, dtype=np.int32); p = np.zeros_like(e)
    lib().orc_edgemaps(C.c_int(np_), ip(e), ip(p))
    return e, p


class Mesh:
    """Oracle mesh (geometry + numbering) from a semtex_amd.mesh dict."""

    def __init__(self, np_, m, svv_mn=-1, svv_eps=0.0, kinvis=1.0):
        self.np = np_
        self._keep = [np.ascontiguousarray(m["vx"], dtype=np.float64),
                      np.ascontiguousarray(m["vy"], dtype=np.float64),
                      np.ascontiguousarray(m["nbr"], dtype=np.int32),
                      np.ascontiguousarray(m["skind"], dtype=np.int32)]
        vx, vy, nbr, sk = self._keep
        self.h = lib().orc_mesh_create(np_, m["nel"], dp(vx), dp(vy), ip(nbr), ip(sk),
                                       svv_mn, svv_eps, kinvis)
        info = np.zeros(10, dtype=np.int32)
        lib().orc_mesh_info(C.c_void_p(self.h), ip(info))
        (self.np, self.nel, self.next, self.nint, self.nglobal, self.nsolve,
         self.nband, self.nbound, self.nPlane, self.planeSize) = [int(v) for v in info]
        self.npnp = self.np * self.np
        self.x = np.zeros(self.nPlane); self.y = np.zeros(self.nPlane)
        lib().orc_mesh_nodes(C.c_void_p(self.h), dp(self.x), dp(self.y))
        self.belmt = np.zeros(self.nbound, dtype=np.int32)
        self.bside = np.zeros(self.nbound, dtype=np.int32)
        lib().orc_mesh_bsides(C.c_void_p(self.h), ip(self.belmt), ip(self.bside))
        self.btog = np.zeros(self.nel * self.next, dtype=np.int32)
        self.bmask = np.zeros_like(self.btog)
        self.emask = np.zeros(self.nel, dtype=np.int32)
        lib().orc_mesh_numbering(C.c_void_p(self.h), ip(self.btog), ip(self.bmask),
                                 ip(self.emask))

    def renumber(self, numbering):
        """Adopt a given numbering (dict with btog, bmask, nglobal, nsolve)."""
        bt = np.ascontiguousarray(numbering["btog"], dtype=np.int32)
        bm = np.ascontiguousarray(numbering["bmask"], dtype=np.int32)
        if lib().orc_mesh_renumber(C.c_void_p(self.h), ip(bt), ip(bm),
                                   int(numbering["nglobal"]), int(numbering["nsolve"])):
            raise ValueError("numbering does not fit the mesh")
        info = np.zeros(10, dtype=np.int32)
        lib().orc_mesh_info(C.c_void_p(self.h), ip(info))
        self.nband = int(info[6])
        lib().orc_mesh_numbering(C.c_void_p(self.h), ip(self.btog), ip(self.bmask),
                                 ip(self.emask))

    def set_cylindrical(self):
        """Switch to cylindrical coordinates (x axial, y radial)."""
        if lib().orc_mesh_set_cylindrical(C.c_void_p(self.h)):
            raise ValueError("a node has y < 0")

    def enumerate(self, optlevel=1):
        """Renumber as semtex's `enumerate -O<optlevel>` (levels 0-3)."""
        if lib().orc_mesh_enumerate(C.c_void_p(self.h), int(optlevel)):
            raise ValueError("unsupported optimisation level")
        info = np.zeros(10, dtype=np.int32)
        lib().orc_mesh_info(C.c_void_p(self.h), ip(info))
        self.nband = int(info[6])
        lib().orc_mesh_numbering(C.c_void_p(self.h), ip(self.btog), ip(self.bmask),
                                 ip(self.emask))

    def factors(self):
        n = self.nPlane
        arrs = [np.zeros(n) for _ in range(8)]
        flags = np.zeros(self.nel * 5, dtype=np.int32)
        lib().orc_mesh_factors(C.c_void_p(self.h), *[dp(a) for a in arrs], ip(flags))
        names = ("Q1", "Q2", "Q3", "Q4", "drdx", "dsdx", "drdy", "dsdy")
        out = dict(zip(names, arrs))
        out["flags"] = flags.reshape(self.nel, 5)
        return out

    def barea(self):
        a = np.zeros(self.nbound * self.np)
        lib().orc_mesh_barea(C.c_void_p(self.h), dp(a))
        return a

    def imass(self):
        a = np.zeros(self.nglobal)
        lib().orc_mesh_imass(C.c_void_p(self.h), dp(a))
        return a

    def __del__(self):
        try:
            lib().orc_mesh_destroy(C.c_void_p(self.h))
        except Exception:
            pass


def field_solve(mesh, nz, lambda2, beta, method, f, bc, u, tol=1e-12,
                step_max=10000, nthreads=0, nrep=1):
    """Field::solve over nz planes; f is destroyed, u updated in place.
    Returns (warnings, mean seconds per all-plane solve, setup excluded)."""
    t = C.c_double(0.0)
    w = lib().orc_field_solve(C.c_void_p(mesh.h), nz, lambda2, beta, method,
                              dp(f), dp(bc), dp(u), tol, step_max, nthreads,
                              nrep, C.byref(t))
    return w, t.value


def field_solve_iters(mesh, nz, lambda2, beta, method, f, bc, u, tol=1e-12,
                      step_max=10000):
    """One Field::solve; returns each plane's PCG iteration count (-1 at
    the step limit, 0 for DIRECT planes and the skipped plane 1)."""
    it = np.zeros(nz, dtype=np.int32)
    lib().orc_field_solve_iters(C.c_void_p(mesh.h), nz, lambda2, beta, method, dp(f),
                                dp(bc), dp(u), tol, step_max,
                                it.ctypes.data_as(C.POINTER(C.c_int)))
    return it


def field_solve_base(mesh, nz, lambda2, beta, base, method, f, bc, u, tol=1e-12,
                     step_max=10000):
    """Field::solve on the planes of modes base..base+nz/2-1; returns each
    plane's PCG iteration count."""
    it = np.zeros(nz, dtype=np.int32)
    lib().orc_field_solve_base(C.c_void_p(mesh.h), nz, lambda2, beta, base, method,
                               dp(f), dp(bc), dp(u), tol, step_max,
                               it.ctypes.data_as(C.POINTER(C.c_int)))
    return it


def dftr(data, tlen, ntrn, sign):
    r = lib().orc_dftr(dp(data), tlen, ntrn, sign)
    if r:
        raise ValueError("dftr: bad length or odd ntrn")
