"""Synthetic, seeded Helmholtz workloads on [0,2pi]^2-type meshes.

Manufactured solutions u* = sum_{m,n<=3} a_mn sin(m x + phi_mn) cos(n y + psi_mn)
with a ~ N(0,1), phi, psi ~ U[0, 2pi) drawn from numpy's PCG64 generator;
for the per-plane equation  lap u - lambda2_k u = f  the forcing is
f = -sum (m^2 + n^2 + lambda2_k) a sin(.) cos(.).
"""
from __future__ import annotations

import numpy as np

from .mesh import rect_mesh


def smooth_coeffs(seed, count=1):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((count, 4, 4))
    phi = rng.uniform(0.0, 2 * np.pi, (count, 4, 4))
    psi = rng.uniform(0.0, 2 * np.pi, (count, 4, 4))
    return a, phi, psi


def smooth_field(x, y, a, phi, psi, lambda2=None):
    """Evaluate u* (and f for lambda2) for coefficient sets of shape (P,4,4)
    at nodes x, y; returns arrays of shape (P, len(x))."""
    P = a.shape[0]
    u = np.zeros((P, x.size))
    f = np.zeros((P, x.size)) if lambda2 is not None else None
    lam = np.broadcast_to(np.asarray(lambda2 if lambda2 is not None else 0.0,
                                     dtype=np.float64), (P,))
    for m in range(4):
        for n in range(4):
            s = np.sin(m * x[None, :] + phi[:, m, n, None]) * np.cos(n * y[None, :] + psi[:, m, n, None])
            u += a[:, m, n, None] * s
            if f is not None:
                f -= (m * m + n * n + lam[:, None]) * a[:, m, n, None] * s
    return u, f


def helmholtz_batch(mesh, nz, lambda2, beta, seed):
    """Per-plane manufactured problems for a field of nz planes:
    returns (f, bc, u_exact) as numpy arrays (nz, planeSize), (nz, nline),
    (nz, planeSize).  Plane k has lambda2_k = lambda2 + (beta*(k>>1))^2;
    the Nyquist plane 1 (nz > 1) is left zero."""
    ps, npl = mesh.planeSize, mesh.nPlane
    f = np.zeros((nz, ps))
    ue = np.zeros((nz, ps))
    planes = [k for k in range(nz) if not (nz > 1 and k == 1)]
    a, phi, psi = smooth_coeffs(seed, len(planes))
    lam = np.array([lambda2 + (beta * (k >> 1)) ** 2 for k in planes])
    u, ff = smooth_field(mesh.x, mesh.y, a, phi, psi, lam)
    for i, k in enumerate(planes):
        ue[k, :npl] = u[i]
        f[k, :npl] = ff[i]
    bc = mesh.bc_lines(ue[:, :npl])
    return f, np.ascontiguousarray(bc), ue


WORKLOADS = {
    # name: (np, K, nz, periodic) -- K x K uniform quads on [0, 2 pi]^2
    "square8_np8": (8, 8, 1, False),
    "square16_np9_nz64": (9, 16, 64, False),
    "square32_np11_nz128": (11, 32, 128, False),
    "periodic24_np9_nz96": (9, 24, 96, True),
}


def workload_mesh(name):
    """Mesh of a named benchmark workload: (np, mesh dict, nz).  All are
    K x K uniform quads on [0, 2 pi]^2, Dirichlet on every edge unless
    periodic both ways."""
    np_, K, nz, per = WORKLOADS[name]
    two_pi = 2 * np.pi
    return np_, rect_mesh(K, K, 0, two_pi, 0, two_pi, periodic_x=per, periodic_y=per), nz
