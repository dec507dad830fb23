"""ctypes prototypes of the semx C ABI (include/semx.h).

Loading fails loudly when libsemx.so has not been built: there is no
fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libsemx.so")

D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int)
VP = C.c_void_p

# name: (restype, argtypes)
PROTOTYPES = {
    "semx_error_string": (C.c_char_p, [C.c_int]),
    "semx_device_init": (C.c_int, [C.c_int]),
    "semx_mesh_create": (C.c_int, [C.c_int, C.c_int, D, D, I, I, C.c_int, C.c_double,
                                   C.c_double, C.POINTER(VP)]),
    "semx_mesh_set_cylindrical": (C.c_int, [VP]),
    "semx_mesh_create_enumerated": (C.c_int, [C.c_int, C.c_int, D, D, I, I, C.c_int, C.c_double,
                                              C.c_double, C.c_int, C.POINTER(VP)]),
    "semx_mesh_create_numbered": (C.c_int, [C.c_int, C.c_int, D, D, I, I, C.c_int, C.c_double,
                                          C.c_double, I, I, C.c_int, C.c_int,
                                          C.POINTER(VP)]),
    "semx_mesh_destroy": (C.c_int, [VP]),
    "semx_mesh_info": (C.c_int, [VP, I]),
    "semx_mesh_nodes": (C.c_int, [VP, D, D]),
    "semx_mesh_bsides": (C.c_int, [VP, I, I]),
    "semx_mesh_numbering": (C.c_int, [VP, I, I, I]),
    "semx_mesh_nsets": (C.c_int, [VP, I]),
    "semx_helm_create": (C.c_int, [VP, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                   C.c_int, C.POINTER(VP)]),
    "semx_helm_create_svv": (C.c_int, [VP, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                       C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                       C.POINTER(VP)]),
    "semx_helm_destroy": (C.c_int, [VP]),
    "semx_helm_sweep_info": (C.c_int, [VP, I]),
    "semx_helm_info": (C.c_int, [VP, I]),
    "semx_helm_element": (C.c_int, [VP, C.c_int, C.c_int, D, D, D]),
    "semx_helm_solve": (C.c_int, [VP, VP, VP, VP, C.c_double, C.c_int, I, VP]),
    "semx_helm_check": (C.c_int, [VP, VP]),
    "semx_helm_iters": (C.c_int, [VP, I, VP]),
    "semx_helm_set_pcg": (C.c_int, [VP, C.c_int, C.c_double, C.c_int]),
    "semx_helm_solve_fields": (C.c_int, [VP, C.c_int, C.POINTER(VP), C.POINTER(VP),
                                         C.POINTER(VP), C.c_double, C.c_int, I, VP]),
    "semx_batch_create": (C.c_int, [C.c_int, C.POINTER(VP), C.POINTER(VP)]),
    "semx_batch_destroy": (C.c_int, [VP]),
    "semx_batch_info": (C.c_int, [VP, I, C.POINTER(C.c_double)]),
    "semx_batch_solve": (C.c_int, [VP, C.POINTER(VP), C.POINTER(VP), C.POINTER(VP),
                                   C.c_double, C.c_int, I, VP]),
    "semx_dft_r": (C.c_int, [VP, C.c_int, C.c_int, C.c_int, VP]),
    "semx_dft_r_noalloc": (C.c_int, [VP, C.c_int, C.c_int, C.c_int, VP]),
    "semx_field_write": (C.c_int, [C.c_char_p, C.c_int, VP, C.c_int, C.c_int, C.c_char_p,
                                   C.POINTER(VP), C.c_char_p, C.c_int, C.c_double, C.c_double,
                                   C.c_double, C.c_double, VP]),
    "semx_field_read": (C.c_int, [C.c_char_p, C.c_int, VP, C.c_int, C.c_int, C.c_char_p,
                                  C.POINTER(VP), I, C.POINTER(C.c_double), C.c_char_p, VP]),
    "semx_gradient": (C.c_int, [VP, C.c_int, C.c_double, VP, C.c_int, VP]),
    "semx_smooth": (C.c_int, [VP, C.c_int, VP, VP]),
    "semx_dns_set_bc": (C.c_int, [VP, VP, VP, VP, VP]),
    "semx_dns_set_iterative": (C.c_int, [VP, C.c_int]),
    "semx_couple": (C.c_int, [VP, VP, C.c_int, C.c_int, C.c_int, C.c_int, VP]),
    "semx_helm_create_field": (C.c_int, [VP, C.c_double, C.c_double, C.c_int, C.c_int,
                                         C.c_int, C.c_int, C.c_int, C.POINTER(VP)]),
    "semx_dns_create": (C.c_int, [VP, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                  C.c_int, C.c_double, C.c_int, C.POINTER(VP)]),
    "semx_dns_create_bc": (C.c_int, [VP, VP, C.c_int, C.c_double, C.c_double, C.c_double,
                                     C.c_int, C.c_int, C.c_double, C.c_int, VP, VP, VP,
                                     C.POINTER(VP)]),
    "semx_dns_destroy": (C.c_int, [VP]),
    "semx_dns_run": (C.c_int, [VP, C.c_int, VP, VP, VP, VP, VP]),
    "semx_dns_info": (C.c_int, [VP, I]),
    "semx_dns_set_graph": (C.c_int, [VP, C.c_int]),
    "semx_dns_graph_info": (C.c_int, [VP, C.POINTER(C.c_int), C.POINTER(C.c_longlong)]),
    "semx_malloc": (C.c_int, [C.POINTER(VP), C.c_size_t]),
    "semx_free": (C.c_int, [VP]),
    "semx_memcpy_htod": (C.c_int, [VP, VP, C.c_size_t]),
    "semx_memcpy_dtoh": (C.c_int, [VP, VP, C.c_size_t]),
    "semx_synchronize": (C.c_int, []),
    "semx_profile_enable": (C.c_int, [C.c_int]),
    "semx_profile_reset": (C.c_int, []),
    "semx_profile_query": (C.c_int, [C.c_char_p, C.POINTER(C.c_double),
                                     C.POINTER(C.c_longlong)]),
}

_lib = None


class SemxError(RuntimeError):
    def __init__(self, status, where=""):
        self.status = status
        msg = load().semx_error_string(status).decode() if _lib else str(status)
        super().__init__(f"{where}: semx status {status} ({msg})")


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise ImportError(
                f"{LIBPATH} is missing: build it with `python -m semtex_amd.build` "
                "(the HIP extension is required; there is no CPU fallback)")
        L = C.CDLL(LIBPATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status, where=""):
    if status != 0:
        raise SemxError(status, where)
