"""Shared set-up for the GPU solve parity tests: meshes, inputs, the two
solvers (GPU through the C ABI, oracle on the CPU) and the error measure."""
import numpy as np

import oracle_lib as O
import semtex_amd as sx
from semtex_amd.mesh import rect_mesh
from semtex_amd.workloads import helmholtz_batch

PERIODIC = "P"


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def solved(P, u):
    """The planes Field::solve writes: all but the Nyquist plane 1
    (src/field.cpp:566), without the padding of each plane."""
    nz = u.shape[0]
    keep = [k for k in range(nz) if not (nz > 1 and k == 1)]
    return u[keep, :P.nPlane]


def natural_lines(P, nz, bcl):
    """Replace natural-side BC values by a smooth synthetic flux (the same
    numbers go to both solvers)."""
    out = bcl.copy()
    np_ = P.np
    rng = np.random.default_rng(5)
    for b in range(P.nbound):
        e, s = P.belmt[b], P.bside[b]
        if s in (1, 3):   # natural sides of these meshes are right/left
            out[:, b * np_:(b + 1) * np_] = rng.standard_normal((nz, 1)) * np.cos(
                P.y[e * np_ * np_ + np.array(sx.side_nodes(np_, s))])[None, :]
    return np.ascontiguousarray(out)


def oracle_numbering(Q):
    """The oracle mesh's numbering in the form semx_mesh_create_numbered
    takes (as semtex's session.num would give it)."""
    return dict(btog=Q.btog, bmask=Q.bmask, nglobal=Q.nglobal, nsolve=Q.nsolve)


def make_case(nx, np_, nz, lam2, beta, dist, bc, seed=11, same_numbering=False):
    """Mesh pair (product, oracle) and seeded forcing / BC lines / exact
    solution for lap u - (lambda2 + (beta k)^2) u = f on every plane.
    same_numbering: the product mesh takes the oracle's numbering (so a
    singular mode 0 is pinned at the same node)."""
    per = bc == PERIODIC
    # periodic meshes span [0, 2 pi]^2 so that the manufactured forcing is
    # periodic and the singular k = 0 system consistent (zero-mean f)
    x1 = 2 * np.pi if (nx >= 16 or per) else 1.0
    m = rect_mesh(nx, nx, 0, x1, 0, x1, bc=("D",) * 4 if per else bc, distort=dist,
                  periodic_x=per, periodic_y=per)
    Q = O.Mesh(np_, m)
    P = sx.Mesh(np_, m, numbering=oracle_numbering(Q) if same_numbering else None)
    f, bcl, ue = helmholtz_batch(P, nz, lam2, beta, seed=seed)
    if not per and "N" in bc:
        bcl = natural_lines(P, nz, bcl)
    return P, Q, f, bcl, ue


def gpu_solve(torch, P, S, f, bc, u0=None, tol=1e-12, step_max=10000):
    """Returns (u, largest PCG iteration count)."""
    dev = torch.device("cuda:0")
    tf = torch.from_numpy(np.ascontiguousarray(f).ravel().copy()).to(dev)
    tb = torch.from_numpy(np.ascontiguousarray(bc).ravel().copy()).to(dev)
    tu = (torch.zeros_like(tf) if u0 is None
          else torch.from_numpy(np.ascontiguousarray(u0).ravel().copy()).to(dev))
    it = sx.solve(S, tf, tb, tu, tol_rel=tol, step_max=step_max)
    torch.cuda.synchronize()
    return tu.cpu().numpy().reshape(f.shape), it


def oracle_solve(Q, nz, lam2, beta, f, bc, method=0, u0=None, tol=1e-12, step_max=10000):
    """Returns (u, per-plane PCG iteration counts)."""
    ff = np.ascontiguousarray(f.copy())
    u = np.zeros_like(ff) if u0 is None else np.ascontiguousarray(u0.copy())
    it = O.field_solve_iters(Q, nz, lam2, beta, method, ff, np.ascontiguousarray(bc), u,
                             tol, step_max)
    return u, it


def drop_mode0_constant(P, ug, uo):
    """A singular mode 0 is pinned at the last-numbered node of each
    solver's own numbering (src/matrix.cpp:167-169): the two agree up to a
    constant, removed here from ug (in place).  Returns the constant."""
    d = (ug[0, :P.nPlane] - uo[0, :P.nPlane]).mean()
    ug[0, :P.nPlane] -= d
    return d
