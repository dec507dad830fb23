"""Cylindrical-coordinate Helmholtz cases (CYLINDRICAL = 1: x axial, y
radial, z azimuthal), shared by the oracle (CPU) and GPU parity tests.

- poisson2 (mesh/poisson2, regress/poisson2.ok): the axisymmetric Poisson
  problem lap c = -16 in a cylinder of radius 0.5 and length 5, 10 x 6
  elements of np = 11 (radii 0, .15, .275, .375, .44, .48, .5), c = 1 - 4y^2
  on the wall, ends and outer radius, the axis BC <A> on y = 0 (for mode 0
  a zero-flux natural condition); semtex's `compare` finds the error at
  noise level (|c - (1 - 4y^2)|_inf < 1e-12).  The forcing is
  premultiplied by y as elliptic/helmholtz.cpp:45-63 does.
- laplace9 (mesh/laplace9, no .ok file): Laplace's equation in the annulus
  0.1 <= y <= 1, 0 <= x <= 2 pi, 2 x 2 elements of np = 11, N_Z = 32,
  essential c = exp(sqrt2 y cos z) sin(y sin z) sin x everywhere, which is
  also the exact (harmonic) solution.
- a solid cylinder 0 <= y <= 1 with the axis, the harmonic
  c = exp(sqrt2 y cos z) cos(y sin z) sin x (mode 0 nonzero on the axis),
  N_Z = 32 (the z content decays as 1.4^k / k!, so 16 planes alias at 1e-6):
  mode 0 is solved with the axis natural (numbering 'c'), modes 1.. with
  it essential zero (numbering 'C'), src/bcmgr.cpp:133-184.
"""
import numpy as np

from semtex_amd.mesh import rect_mesh
from semtex_amd.workloads import fourier_pack


def graded_mesh(xs, ys, bc):
    """Quads on the tensor grid of breakpoints xs x ys; bc: kinds of the
    bottom, right, top and left sides."""
    xs, ys = np.asarray(xs, float), np.asarray(ys, float)
    nx, ny = len(xs) - 1, len(ys) - 1
    m = rect_mesh(nx, ny, 0.0, 1.0, 0.0, 1.0, bc=bc)
    m["vx"] = xs[np.rint(np.asarray(m["vx"]) * nx).astype(int)]
    m["vy"] = ys[np.rint(np.asarray(m["vy"]) * ny).astype(int)]
    return m


POISSON2_Y = [0.0, 0.15, 0.275, 0.375, 0.44, 0.48, 0.5]


def poisson2_mesh():
    return graded_mesh(np.linspace(0.0, 5.0, 11), POISSON2_Y, ("N", "D", "D", "D"))


def poisson2_problem(M):
    """(f premultiplied by y, BC lines, exact) on a (1, planeSize) plane."""
    from semtex_amd.mesh import bc_lines
    f = np.zeros((1, M.planeSize))
    f[0, :M.nPlane] = -16.0 * M.y
    ex = np.zeros((1, M.planeSize))
    ex[0, :M.nPlane] = 1.0 - 4.0 * M.y ** 2
    bc = bc_lines(ex[:, :M.nPlane], M.belmt, M.bside, M.np)
    for b in range(M.nbound):
        if M.bside[b] == 0:                       # the axis: zero flux
            bc[:, b * M.np:(b + 1) * M.np] = 0.0
    return f, np.ascontiguousarray(bc), ex


def laplace9_mesh():
    return graded_mesh([0.0, np.pi, 2 * np.pi], [0.1, 0.5, 1.0], ("D", "D", "D", "D"))


def laplace9_exact(x, y, z):
    return np.exp(np.sqrt(2.0) * y * np.cos(z)) * np.sin(y * np.sin(z)) * np.sin(x)


def tube_meshes():
    """(axis natural, axis essential): the same 2 x 3 elements of a solid
    cylinder, radii 0, 0.3, 0.7, 1."""
    xs, ys = [0.0, np.pi, 2 * np.pi], [0.0, 0.3, 0.7, 1.0]
    return (graded_mesh(xs, ys, ("N", "D", "D", "D")),
            graded_mesh(xs, ys, ("D", "D", "D", "D")))


def tube_exact(x, y, z):
    return np.exp(np.sqrt(2.0) * y * np.cos(z)) * np.cos(y * np.sin(z)) * np.sin(x)


def fourier_exact(M, nz, fn):
    """(nz, planeSize) Fourier planes of fn on the mesh nodes (Nyquist plane
    zeroed) and its BC lines (nz, nbound*np)."""
    from semtex_amd.mesh import bc_lines
    z = 2 * np.pi * np.arange(nz) / nz
    phys = np.stack([fn(M.x, M.y, zk) for zk in z])
    F = np.zeros((nz, M.planeSize))
    F[:, :M.nPlane] = fourier_pack(phys) if nz > 1 else phys
    if nz > 1:
        F[1] = 0.0
    bc = bc_lines(F[:, :M.nPlane], M.belmt, M.bside, M.np)
    return F, np.ascontiguousarray(bc)


def axis_lines(M, bc, mode0):
    """BC lines of the axis sides (side 0 on y = 0): zero flux for mode 0,
    zero value for the other modes -- both zero."""
    out = bc.copy()
    for b in range(M.nbound):
        if M.bside[b] == 0:
            out[:, b * M.np:(b + 1) * M.np] = 0.0
    return out


def laplace9_flux_left(x, y, z):
    """-d/dx of laplace9_exact: the outward flux through the end x = 0."""
    return -np.exp(np.sqrt(2.0) * y * np.cos(z)) * np.sin(y * np.sin(z)) * np.cos(x)


def annulus_natural_mesh():
    """laplace9's annulus with the end x = 0 (left side, 0.1 <= y <= 1)
    natural."""
    return graded_mesh([0.0, np.pi, 2 * np.pi], [0.1, 0.5, 1.0], ("D", "D", "D", "N"))


def natural_lines(M, nz, bc, fn_flux, side):
    """bc with the lines of `side` replaced by the Fourier coefficients of
    the flux fn_flux there."""
    F, fl = fourier_exact(M, nz, fn_flux)
    out = bc.copy()
    for b in range(M.nbound):
        if M.bside[b] == side:
            out[:, b * M.np:(b + 1) * M.np] = fl[:, b * M.np:(b + 1) * M.np]
    return out


def swirl_annulus_mesh():
    """0 <= x <= 1, 0.5 <= y <= 1.5, 2 x 2 elements, all sides essential."""
    return graded_mesh([0.0, 0.5, 1.0], [0.5, 1.0, 1.5], ("D", "D", "D", "D"))


def coupled_modal_exact(M, nz):
    """Fourier planes (nz, planeSize) of v and w whose coupled fields are,
    in every (x, y) plane, v~ = y^(k+1) and w~ = y^(k-1) for k = 1 (w~ = 1)
    and k = 2, and ln y for w~ at k = 1 added: solutions of the coupled
    viscous operators d2/dx2 + d2/dy2 + (1/y) d/dy - (k +- 1)^2 / y^2 with
    no forcing.  v = (v~ + w~)/2 is real, w = i (w~ - v~)/2 imaginary."""
    Fv = np.zeros((nz, M.planeSize))
    Fw = np.zeros((nz, M.planeSize))
    y = M.y
    for k in (1, 2):
        vt = y ** (k + 1)
        wt = y ** (k - 1) + (np.log(y) if k == 1 else 0.0)
        Fv[2 * k, :M.nPlane] = 0.5 * (vt + wt)
        Fw[2 * k + 1, :M.nPlane] = 0.5 * (wt - vt)
    return Fv, Fw
