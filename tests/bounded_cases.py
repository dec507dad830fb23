"""DNS cases on meshes with boundaries, shared by the oracle (CPU) and the
GPU parity tests of the computed pressure boundary conditions.

kovas1 (mesh/kovas1, regress/kovas1.ok): Kovasznay flow on four elements
of np = 11 over [-0.5, 1] x [-0.5, 0.5], Re = 40, first-order time
stepping (N_TIME = 1), 400 steps of 0.008, the alternating skew-symmetric
nonlinear terms (dns's default, ADVECTION = 1).  The velocity takes the exact
solution on every side; the pressure takes the computed high-order flux
(<H> p = 0) everywhere, so its mode 0 is singular and pinned at the last
unknown of semtex's numbering -- the regression session is numbered by
`enumerate -O3` (test/testregress), and the run starts from the exact
solution (`compare kovas1 > kovas1.rst`).
"""
import numpy as np

from semtex_amd.mesh import bc_lines, node_mesh, rect_mesh

KOVAS_XY = [(-0.5, -0.5), (0, -0.5), (1, -0.5), (-0.5, 0), (0, 0), (1, 0),
            (-0.5, 0.5), (0, 0.5), (1, 0.5)]
KOVAS_EL = [(0, 1, 4, 3), (1, 2, 5, 4), (3, 4, 7, 6), (4, 5, 8, 7)]
KOVAS = dict(np=11, nstep=400, dt=0.008, Re=40.0, norder=1, advection=1)   # dns defaults: ADVECTION 1
KOVAS_OK = ("8.392e-07", "2.901e-07", "5.012e-01")      # regress/kovas1.ok: u, v, p


def kovas_meshes():
    """(velocity mesh, pressure mesh): the same elements, sides essential
    for the velocity and natural for the pressure."""
    return (node_mesh(KOVAS_XY, KOVAS_EL, default_kind="D"),
            node_mesh(KOVAS_XY, KOVAS_EL, default_kind="N"))


def kovas_exact(x, y, Re=40.0):
    lam = Re / 2 - np.sqrt(0.25 * Re * Re + 4 * np.pi ** 2)
    return (1 - np.exp(lam * x) * np.cos(2 * np.pi * y),
            lam / (2 * np.pi) * np.exp(lam * x) * np.sin(2 * np.pi * y),
            (1 - np.exp(lam * x)) / 2)


def kovas_fields(M):
    """Initial u, v, p (exact, (1, planeSize) each) and velocity BC lines
    (1, nline) on mesh M (anything with x, y, nPlane, planeSize, belmt,
    bside, np)."""
    out = []
    for f in kovas_exact(M.x, M.y):
        a = np.zeros((1, M.planeSize))
        a[0, :M.nPlane] = f
        out.append(a)
    bcu = np.ascontiguousarray(bc_lines(out[0][:, :M.nPlane], M.belmt, M.bside, M.np))
    bcv = np.ascontiguousarray(bc_lines(out[1][:, :M.nPlane], M.belmt, M.bside, M.np))
    return out, (bcu, bcv)


def error_norms(M, u, v, p):
    """compare's norm_inf of u, v, p against the exact solution, as text."""
    ue, ve, pe = kovas_exact(M.x, M.y)
    n = M.nPlane
    return tuple(f"{np.abs(a[0, :n] - e).max():.3e}" for a, e in ((u, ue), (v, ve), (p, pe)))


# --------------------------------------------------------------------------
# A sheared channel with an outflow side, for GPU/oracle parity in 3D:
# elements of a distorted 4 x 3 grid mapped by x' = x + 0.35 y,
# y' = 0.9 y + 0.2 x (slanted boundaries, so no normal is axis-aligned).
# Velocity: bottom, top, left essential, right natural (zero flux);
# pressure: the reverse (computed flux on bottom, top, left; p = 0 on the
# right).  The velocity BC lines come from a divergence-free field A.  A
# run from A itself keeps u.n fixed on the essential sides (the flux's
# time-derivative term cancels); a run from a shifted field B changes u.n
# there over the first N_TIME steps, so that term carries a pressure
# transient of order |B - A| / dt that later decays.
# --------------------------------------------------------------------------

def sheared_meshes(nx=4, ny=3):
    out = []
    for kinds in (("D", "N", "D", "D"), ("N", "D", "N", "N")):
        m = rect_mesh(nx, ny, 0.0, 2.0, -0.5, 0.5, bc=kinds, distort=0.15)
        x, y = m["vx"].copy(), m["vy"].copy()
        m["vx"] = x + 0.35 * y
        m["vy"] = 0.9 * y + 0.2 * x
        out.append(m)
    return out


def _physical(x, y, z, shift):
    """A divergence-free field with z-variation in every component."""
    return (1.0 + 0.2 * np.sin(y + z + shift) + 0.1 * x
            - 0.2 * np.cos(x + y) * np.sin(z + shift),
            0.3 * np.cos(x + shift) + 0.1 * np.cos(2 * z) - 0.1 * y,
            0.2 * np.sin(x + y) * np.cos(z + shift))


def fourier_planes(M, nz, shift, pack):
    """(u, v, w) in Fourier space, (nz, planeSize) each, of the smooth
    fields above; pack maps physical planes (nz, nPlane) to semtex's
    Fourier packing."""
    z = 2 * np.pi * np.arange(nz) / nz
    out = []
    for c in range(3):
        phys = np.stack([_physical(M.x, M.y, zk, shift)[c] for zk in z])
        f = np.zeros((nz, M.planeSize))
        f[:, :M.nPlane] = pack(phys)
        if nz > 1:
            f[1] = 0.0        # the Nyquist plane, which Field::solve never writes
        out.append(f)
    return out


def sheared_bc_lines(M, fields):
    """Velocity BC lines of the Fourier fields: their values on the
    essential sides, zero flux on the natural (right, side 1) ones."""
    lines = []
    for f in fields:
        bl = bc_lines(f[:, :M.nPlane], M.belmt, M.bside, M.np)
        for b in range(M.nbound):
            if M.bside[b] == 1:
                bl[:, b * M.np:(b + 1) * M.np] = 0.0
        lines.append(np.ascontiguousarray(bl))
    return lines


# --------------------------------------------------------------------------
# Time-dependent velocity BCs: the decaying Taylor vortex in the box
# [0,2]^2 (2 x 2 elements, np = 11), velocity essential all round with the
# exact values at each step's end time (BCmgr re-evaluates them there,
# dns/integrate.cpp:196-201), pressure natural all round (computed flux).
# --------------------------------------------------------------------------

BOX = dict(np=11, nu=0.05, dt=0.01, nstep=20, norder=2)


def box_meshes():
    from semtex_amd.mesh import rect_mesh
    return (rect_mesh(2, 2, 0.0, 2.0, 0.0, 2.0, bc=("D", "D", "D", "D")),
            rect_mesh(2, 2, 0.0, 2.0, 0.0, 2.0, bc=("N", "N", "N", "N")))


def box_exact(x, y, t, nu):
    f = np.exp(-2 * np.pi ** 2 * nu * t)
    return (-np.cos(np.pi * x) * np.sin(np.pi * y) * f, np.sin(np.pi * x) * np.cos(np.pi * y) * f)


def box_lines(M, t, nu):
    """(u, v) planes (1, planeSize) at time t and their BC lines."""
    out, lines = [], []
    for a in box_exact(M.x, M.y, t, nu):
        F = np.zeros((1, M.planeSize))
        F[0, :M.nPlane] = a
        out.append(F)
        lines.append(np.ascontiguousarray(bc_lines(F[:, :M.nPlane], M.belmt, M.bside, M.np)))
    return out, lines
