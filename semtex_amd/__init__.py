"""semtex_amd -- MI355X-native spectral-element / Fourier Helmholtz core.

Python mirror of the semtex classes on the hot path, over the C ABI of
libsemx.so (include/semx.h):

  Mesh            Element construction + global numbering
                  (src/element.cpp:30-106, src/bcmgr.cpp:586-824)
  ModalMatrixSys  one Helmholtz system per Fourier mode (src/matrix.cpp:33-110)
  solve()         Field::solve (src/field.cpp:509-725)
  dft_r()         dDFTr along z (femlib/fourier.c:47-171)
  gradient()      AuxField::gradient (src/auxfield.cpp:514-665)
  smooth()        Field::smooth (src/field.cpp:190-236)
  DNS             the time loop of integrate (dns/integrate.cpp:79-221)

Device buffers are torch tensors on the GPU (float64, contiguous); torch is
used only for memory and streams.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi
from ._capi import SemxError, check, load
from .mesh import bc_lines, rect_mesh, side_nodes

DIRECT, JACPCG, MIXED = 0, 1, 2
PCG_FP64, PCG_FP32_REFINED, PCG_BF16_REFINED = 0, 1, 2
E_MAXITER = 3
FORWARD, INVERSE = 1, -1

__all__ = ["Mesh", "ModalMatrixSys", "solve", "solve_fields", "Batch", "dft_r", "gradient", "smooth", "DNS",
           "DIRECT", "JACPCG", "MIXED",
           "PCG_FP64", "PCG_FP32_REFINED", "PCG_BF16_REFINED",
           "rect_mesh", "bc_lines", "side_nodes", "SemxError", "load"]


def _dp(a):
    return a.ctypes.data_as(_capi.D)


def _ip(a):
    return a.ctypes.data_as(_capi.I)


class Mesh:
    """Elements + numbering of a quad mesh (see semtex_amd.mesh)."""

    def __init__(self, np_, m, svv_mn=-1, svv_eps=0.0, kinvis=1.0, numbering=None,
                 cylindrical=False):
        """numbering: None for the library's own; semtex's numbering as read
        from a session.num file (semtex_amd.numfile.read_num: a dict with
        btog, bmask, nglobal, nsolve); or ("enumerate", level) to number as
        semtex's `enumerate -O<level>` would (levels 0-3).
        cylindrical: the CYLINDRICAL token -- x axial, y radial (y >= 0);
        Helmholtz systems built on the mesh are premultiplied by y, so the
        forcing handed to a solve must be y f (AuxField::mulY, as
        elliptic/helmholtz.cpp does)."""
        L = load()
        vx = np.ascontiguousarray(m["vx"], dtype=np.float64)
        vy = np.ascontiguousarray(m["vy"], dtype=np.float64)
        nbr = np.ascontiguousarray(m["nbr"], dtype=np.int32)
        sk = np.ascontiguousarray(m["skind"], dtype=np.int32)
        h = C.c_void_p()
        if numbering is None:
            check(L.semx_mesh_create(np_, int(m["nel"]), _dp(vx), _dp(vy), _ip(nbr), _ip(sk),
                                     svv_mn, svv_eps, kinvis, C.byref(h)), "semx_mesh_create")
        elif isinstance(numbering, tuple) and numbering[0] == "enumerate":
            check(L.semx_mesh_create_enumerated(np_, int(m["nel"]), _dp(vx), _dp(vy), _ip(nbr),
                                                _ip(sk), svv_mn, svv_eps, kinvis,
                                                int(numbering[1]), C.byref(h)),
                  "semx_mesh_create_enumerated")
        else:
            bt = np.ascontiguousarray(numbering["btog"], dtype=np.int32)
            bm = np.ascontiguousarray(numbering["bmask"], dtype=np.int32)
            check(L.semx_mesh_create_numbered(np_, int(m["nel"]), _dp(vx), _dp(vy), _ip(nbr),
                                              _ip(sk), svv_mn, svv_eps, kinvis, _ip(bt),
                                              _ip(bm), int(numbering["nglobal"]),
                                              int(numbering["nsolve"]), C.byref(h)),
                  "semx_mesh_create_numbered")
        self.h = h
        self.cylindrical = bool(cylindrical)
        if cylindrical:
            check(L.semx_mesh_set_cylindrical(h), "semx_mesh_set_cylindrical")
        info = np.zeros(10, dtype=np.int32)
        check(L.semx_mesh_info(h, _ip(info)))
        (self.np, self.nel, self.next, self.nint, self.nglobal, self.nsolve,
         self.nband, self.nbound, self.nPlane, self.planeSize) = [int(v) for v in info]
        self.npnp = self.np * self.np
        self.x = np.zeros(self.nPlane)
        self.y = np.zeros(self.nPlane)
        check(L.semx_mesh_nodes(h, _dp(self.x), _dp(self.y)))
        self.belmt = np.zeros(max(self.nbound, 1), dtype=np.int32)
        self.bside = np.zeros(max(self.nbound, 1), dtype=np.int32)
        check(L.semx_mesh_bsides(h, _ip(self.belmt), _ip(self.bside)))
        self.belmt = self.belmt[:self.nbound]
        self.bside = self.bside[:self.nbound]
        self.btog = np.zeros(self.nel * self.next, dtype=np.int32)
        self.bmask = np.zeros_like(self.btog)
        self.emask = np.zeros(self.nel, dtype=np.int32)
        check(L.semx_mesh_numbering(h, _ip(self.btog), _ip(self.bmask), _ip(self.emask)))
        ns = C.c_int()
        check(L.semx_mesh_nsets(h, C.byref(ns)))
        self.nsets = ns.value
        self.nline = self.nbound * self.np

    def bc_lines(self, field):
        return bc_lines(field, self.belmt, self.bside, self.np)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                load().semx_mesh_destroy(self.h)
                self.h = None
        except Exception:
            pass


class ModalMatrixSys:
    """Helmholtz systems for modes base..base+nmodes-1 of a field of nz
    planes (src/matrix.cpp:33-110)."""

    def __init__(self, mesh, lambda2, beta, base_mode, nmodes, nz, method=DIRECT,
                 svv_mz=-1, svv_epsz=0.0, kinvis=1.0, nmode_total=None, field="c"):
        """svv_mz >= 0: semtex's Fourier-direction SVV (SVV_MZ, SVV_EPSZ,
        KINVIS) on the betak2 of modes above svv_mz; nmode_total is the
        field's Fourier mode count (default base_mode + nmodes).  field:
        the field's name, which picks the Fourier constant on cylindrical
        meshes (Field::modeConstant: 'v' / 'w' are the coupled v~ / w~)."""
        L = load()
        h = C.c_void_p()
        if field != "c" and svv_mz < 0:
            check(L.semx_helm_create_field(mesh.h, float(lambda2), float(beta), int(base_mode),
                                           int(nmodes), int(nz), int(method), ord(field),
                                           C.byref(h)), "semx_helm_create_field")
        elif svv_mz >= 0:
            nmt = base_mode + nmodes if nmode_total is None else nmode_total
            check(L.semx_helm_create_svv(mesh.h, float(lambda2), float(beta), int(base_mode),
                                         int(nmodes), int(nz), int(method), int(svv_mz),
                                         float(svv_epsz), float(kinvis), int(nmt),
                                         C.byref(h)), "semx_helm_create_svv")
        else:
            check(L.semx_helm_create(mesh.h, float(lambda2), float(beta), int(base_mode),
                                     int(nmodes), int(nz), int(method), C.byref(h)),
                  "semx_helm_create")
        self.h = h
        self.mesh = mesh
        self.nz = nz
        self.nmodes = nmodes
        info = np.zeros(8, dtype=np.int32)
        check(L.semx_helm_info(h, _ip(info)))
        self.info = dict(nmodes=int(info[0]), nz=int(info[1]), method=int(info[2]),
                         nsolve0=int(info[3]), ldab=int(info[4]), nb=int(info[5]),
                         factor_kB=int(info[6]), singular0=int(info[7]))

    def set_pcg(self, precision=PCG_FP64, inner_tol=1e-5, max_outer=12):
        """Precision of the JACPCG planes (FP64, or FP32 inner PCG -- element
        operator in FP32 or on the BF16 matrix cores -- under FP64 iterative
        refinement)."""
        check(load().semx_helm_set_pcg(self.h, int(precision), float(inner_tol),
                                       int(max_outer)), "semx_helm_set_pcg")

    def plane_iters(self, stream=None):
        """Per-plane JACPCG iteration counts of the last solve (0 for planes
        solved directly and for the Nyquist plane; negative at the limit)."""
        it = np.zeros(self.nz, dtype=np.int32)
        check(load().semx_helm_iters(self.h, _ip(it), _stream(stream)), "semx_helm_iters")
        return it

    def element(self, m, e):
        M = self.mesh
        hbb = np.zeros((M.next, M.next))
        hbi = np.zeros((M.next, max(M.nint, 1)))
        hii = np.zeros((max(M.nint, 1), max(M.nint, 1)))
        check(load().semx_helm_element(self.h, m, e, _dp(hbb), _dp(hbi), _dp(hii)))
        return hbb, hbi[:, :M.nint], hii[:M.nint, :M.nint]

    def sweep_info(self):
        """Band-sweep and PCG layout: dict(npart, pick={1: , 2: , 3: },
        coop_w, pcg_w)."""
        info = np.zeros(6, dtype=np.int32)
        check(load().semx_helm_sweep_info(self.h, _ip(info)), "semx_helm_sweep_info")
        return dict(npart=int(info[0]), pick={f: int(info[f]) for f in (1, 2, 3)},
                    coop_w=int(info[4]), pcg_w=int(info[5]))

    def __del__(self):
        try:
            if getattr(self, "h", None):
                load().semx_helm_destroy(self.h)
                self.h = None
        except Exception:
            pass


def _ptr(t):
    if t is None:
        return None
    import torch
    assert t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()
    return C.c_void_p(t.data_ptr())


def solve(msys, f, bc, u, tol_rel=1e-12, step_max=10000, stream=None, iters=True):
    """Field::solve over all planes: f (destroyed), bc, u are device tensors
    of nz*planeSize, nz*nbound*np and nz*planeSize doubles.  With iters the
    call waits for the stream and returns the largest PCG iteration count
    (0 for DIRECT); a plane at its step limit gives a RuntimeWarning, as
    semtex's "step limit exceeded" WARNING (src/field.cpp:713)."""
    import torch
    M = msys.mesh
    assert f.numel() == msys.nz * M.planeSize and u.numel() == msys.nz * M.planeSize
    if M.nbound:
        assert bc is not None and bc.numel() == msys.nz * M.nline
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    it = C.c_int(0)
    st = load().semx_helm_solve(msys.h, _ptr(f), _ptr(bc), _ptr(u), float(tol_rel),
                                int(step_max), C.byref(it) if iters else None,
                                C.c_void_p(stream))
    if st == E_MAXITER:
        import warnings
        warnings.warn("Field::solve: step limit exceeded", RuntimeWarning)
    else:
        check(st, "semx_helm_solve")
    return it.value


def helm_check(msys, stream=None):
    """Wait for the stream and raise SemxError(SEMX_E_TIMEOUT) if any solve
    of this system since the last check had a cooperative band sweep whose
    bounded hand-off wait expired (semx_helm_check)."""
    check(load().semx_helm_check(msys.h, _stream(stream)), "semx_helm_check")


def solve_fields(msys, fs, bcs, us, tol_rel=1e-12, step_max=10000, stream=None, iters=True):
    """Field::solve for up to 3 fields sharing msys (as u, v, w do): one
    sweep of each mode's factor for all of them.  Lists of device tensors;
    bcs may be None on a mesh without boundaries."""
    import torch
    M = msys.mesh
    nf = len(fs)
    assert len(us) == nf and 1 <= nf <= 3
    for t in list(fs) + list(us):
        assert t.numel() == msys.nz * M.planeSize
    arr = C.c_void_p * nf
    fa = arr(*[_ptr(t) for t in fs])
    ua = arr(*[_ptr(t) for t in us])
    ba = arr(*[_ptr(t) for t in bcs]) if bcs is not None else None
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    it = C.c_int(0)
    st = load().semx_helm_solve_fields(msys.h, nf, fa, ba, ua, float(tol_rel), int(step_max),
                                       C.byref(it) if iters else None, C.c_void_p(stream))
    if st == E_MAXITER:
        import warnings
        warnings.warn("Field::solve: step limit exceeded", RuntimeWarning)
    else:
        check(st, "semx_helm_solve_fields")
    return it.value


class Batch:
    """Independent 2D JACPCG problems (systems made with nz=1, nmodes=1,
    method JACPCG, equal set_pcg settings) solved in one launch, one
    workgroup per problem."""

    def __init__(self, systems):
        self.systems = list(systems)
        arr = (C.c_void_p * len(self.systems))(*[S.h for S in self.systems])
        h = C.c_void_p()
        check(load().semx_batch_create(len(self.systems), arr, C.byref(h)), "semx_batch_create")
        self.h = h

    def solve(self, fs, bcs, us, tol_rel=1e-10, step_max=10000, stream=None, iters=True):
        n = len(self.systems)
        assert len(fs) == n and len(us) == n and (bcs is None or len(bcs) == n)
        arr = C.c_void_p * n
        fa = arr(*[_ptr(t) for t in fs])
        ua = arr(*[_ptr(t) for t in us])
        ba = arr(*[_ptr(t) if t is not None else None for t in bcs]) if bcs is not None else None
        it = C.c_int(0)
        st = load().semx_batch_solve(self.h, fa, ba, ua, float(tol_rel), int(step_max),
                                     C.byref(it) if iters else None, _stream(stream))
        if st == E_MAXITER:
            import warnings
            warnings.warn("Field::solve: step limit exceeded", RuntimeWarning)
        else:
            check(st, "semx_batch_solve")
        return it.value

    OPERATORS = ("block", "wave", "mfma")

    def info(self):
        """(element operator per problem -- 'block', 'wave' or 'mfma', None
        before the first solve -- and the tuner's host time in ms)."""
        n = len(self.systems)
        opk = np.zeros(n, dtype=np.int32)
        ms = C.c_double(0.0)
        check(load().semx_batch_info(self.h, _ip(opk), C.byref(ms)), "semx_batch_info")
        return [self.OPERATORS[k] if k >= 0 else None for k in opk], ms.value

    def __del__(self):
        try:
            if getattr(self, "h", None):
                load().semx_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass


def _stream(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream().cuda_stream
    return C.c_void_p(stream)


def couple(v, w, nz, n, base_mode=0, direction=1, stream=None):
    """AuxField::couple / Field::coupleBCs on device arrays (in place):
    direction 1 forms v~ = v + i w, w~ = v - i w per mode, -1 undoes it."""
    check(load().semx_couple(C.c_void_p(v.data_ptr()), C.c_void_p(w.data_ptr()), int(nz),
                             int(n), int(base_mode), int(direction), _stream(stream)),
          "semx_couple")


def dft_r(data, tlen, ntrn, sign, stream=None):
    """dDFTr in place on a device tensor of tlen*ntrn doubles."""
    assert data.numel() >= tlen * ntrn
    check(load().semx_dft_r(_ptr(data), int(tlen), int(ntrn), int(sign), _stream(stream)),
          "semx_dft_r")


def dft_r_noalloc(data, tlen, ntrn, sign, stream=None):
    """dDFTr in place through the allocation-free entry point (static device
    scratch; tlen <= 256 and tlen*ntrn <= SEMX_DFT_NOALLOC_MAX)."""
    assert data.numel() >= tlen * ntrn
    check(load().semx_dft_r_noalloc(_ptr(data), int(tlen), int(ntrn), int(sign),
                                    _stream(stream)), "semx_dft_r_noalloc")


def write_field(path, mesh, nz, fields, session="session", step=0, time=0.0, dt=0.0,
                kinvis=0.0, beta=1.0, append=False, stream=None):
    """Write device fields (Fourier space) as one dump of a semtex field
    file; fields: {name letter: tensor of nz*planeSize}."""
    names = "".join(fields)
    arr = (C.c_void_p * len(fields))(*[_ptr(t) for t in fields.values()])
    for t in fields.values():
        assert t.numel() == nz * mesh.planeSize
    check(load().semx_field_write(path.encode(), int(bool(append)), mesh.h, int(nz), len(fields),
                                  names.encode(), arr, session.encode(), int(step), float(time),
                                  float(dt), float(kinvis), float(beta), _stream(stream)),
          "semx_field_write")


def read_field(path, mesh, nz, fields, dump=0, stream=None):
    """Read dump `dump` of a semtex field file into device fields (Fourier
    space), {name letter: tensor}; returns (step, time, field names in the
    file)."""
    names = "".join(fields)
    arr = (C.c_void_p * max(len(fields), 1))(*[_ptr(t) for t in fields.values()])
    step = C.c_int(0)
    t = C.c_double(0.0)
    buf = C.create_string_buffer(64)
    check(load().semx_field_read(path.encode(), int(dump), mesh.h, int(nz), len(fields),
                                 names.encode(), arr, C.byref(step), C.byref(t), buf,
                                 _stream(stream)), "semx_field_read")
    return step.value, t.value, buf.value.decode()


def gradient(mesh, nz, beta, data, direction, stream=None):
    """AuxField::gradient in place (direction 0: x, 1: y, 2: z in Fourier space)."""
    assert data.numel() == nz * mesh.planeSize
    check(load().semx_gradient(mesh.h, int(nz), float(beta), _ptr(data), int(direction),
                               _stream(stream)), "semx_gradient")


def smooth(mesh, nz, data, stream=None):
    """Field::smooth in place: mass-weighted average along element edges."""
    assert data.numel() == nz * mesh.planeSize
    check(load().semx_smooth(mesh.h, int(nz), _ptr(data), _stream(stream)), "semx_smooth")


class DNS:
    """Device time stepping of semtex's dns solver (integrate,
    dns/integrate.cpp:79-221); fields in Fourier space.

    On a periodic mesh, pressure_mesh and bc are left out.  On a mesh with
    boundaries, mesh numbers the velocity and pressure_mesh the pressure
    (two numberings of one element mesh, e.g. from session.num), and bc is
    the velocity BC lines (u, v[, w]): device tensors of nz*nbound*np in
    Fourier space, fixed in time; the pressure takes the computed
    high-order flux on its natural sides."""

    def __init__(self, mesh, nz, beta, dt, kinvis, norder=2, advection=0, tol_rel=1e-12,
                 step_max=500, pressure_mesh=None, bc=None):
        h = C.c_void_p()
        if pressure_mesh is None:
            check(load().semx_dns_create(mesh.h, int(nz), float(beta), float(dt),
                                         float(kinvis), int(norder), int(advection),
                                         float(tol_rel), int(step_max), C.byref(h)),
                  "semx_dns_create")
        else:
            ncom = 3 if nz > 1 else 2
            if bc is None or len(bc) < ncom:
                raise ValueError(f"bc needs {ncom} velocity BC line sets")
            nl = nz * mesh.nbound * mesh.np
            for t in bc[:ncom]:
                assert t.numel() == nl
            check(load().semx_dns_create_bc(mesh.h, pressure_mesh.h, int(nz), float(beta),
                                            float(dt), float(kinvis), int(norder),
                                            int(advection), float(tol_rel), int(step_max),
                                            _ptr(bc[0]), _ptr(bc[1]),
                                            _ptr(bc[2]) if ncom == 3 else None, C.byref(h)),
                  "semx_dns_create_bc")
        self.h = h
        self.mesh = mesh
        self.pressure_mesh = pressure_mesh
        self.nz = nz

    def run(self, nstep, u, v, w, p, stream=None):
        n = self.nz * self.mesh.planeSize
        for t in (u, v, p) + ((w,) if self.nz > 1 else ()):
            assert t.numel() == n
        check(load().semx_dns_run(self.h, int(nstep), _ptr(u), _ptr(v),
                                  _ptr(w) if self.nz > 1 else None, _ptr(p),
                                  _stream(stream)), "semx_dns_run")

    def set_bc(self, lines, stream=None):
        """New velocity BC lines (device tensors, nz*nline each) for the
        steps that follow: time-dependent BC values (semx_dns_set_bc)."""
        ptrs = [_ptr(t) for t in lines] + [None] * (3 - len(lines))
        check(load().semx_dns_set_bc(self.h, ptrs[0], ptrs[1], ptrs[2], _stream(stream)),
              "semx_dns_set_bc")

    def set_iterative(self, level):
        """The ITERATIVE token (dns -i): 1 solves the velocity with JACPCG,
        2 also the pressure with MIXED, 0 with DIRECT (before the first
        step; semx_dns_set_iterative)."""
        check(load().semx_dns_set_iterative(self.h, int(level)), "semx_dns_set_iterative")

    @property
    def steps(self):
        info = np.zeros(4, dtype=np.int32)
        check(load().semx_dns_info(self.h, _ip(info)))
        return int(info[0])

    def set_graph(self, on=True):
        """Replay steady time steps from HIP graphs (one per buffer-role
        assignment of the step, captured on first use)."""
        check(load().semx_dns_set_graph(self.h, int(bool(on))), "semx_dns_set_graph")

    def graph_info(self):
        """(graphs captured, steps replayed from graphs)."""
        ng = C.c_int(0)
        rp = C.c_longlong(0)
        check(load().semx_dns_graph_info(self.h, C.byref(ng), C.byref(rp)),
              "semx_dns_graph_info")
        return ng.value, rp.value

    def __del__(self):
        try:
            if getattr(self, "h", None):
                load().semx_dns_destroy(self.h)
                self.h = None
        except Exception:
            pass
