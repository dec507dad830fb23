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


def fourier_pack(phys):
    """z-planes of physical values (nz, n) -> semtex's Fourier planes
    (dDFTr forward packing [Re0, Re_nz/2, Re1, Im1, ...], scaled 1/nz)."""
    nz = phys.shape[0]
    if nz == 1:
        return phys.copy()
    X = np.fft.rfft(phys, axis=0) / nz
    out = np.zeros_like(phys)
    out[0] = X[0].real
    out[1] = X[nz // 2].real
    for j in range(1, nz // 2):
        out[2 * j] = X[j].real
        out[2 * j + 1] = X[j].imag
    return out


def taylor_green_3d(mesh, nz, beta=1.0, seed=3, amp=1e-3):
    """Fourier-space (u, v, w) of the Taylor-Green vortex
    u = sin x cos y cos z, v = -cos x sin y cos z, w = 0 on [0, 2 pi]^3
    (z = 2 pi k / (beta nz)), plus a divergence-free perturbation
    curl(A) with A a seeded random periodic potential of wavenumbers
    |k| <= 3, scaled to max |curl A| = amp.  Arrays (nz, planeSize)."""
    rng = np.random.default_rng(seed)
    x, y = mesh.x, mesh.y
    z = 2 * np.pi * np.arange(nz) / (beta * nz)
    X, Z = x[None, :], z[:, None]
    Y = y[None, :]
    u = np.sin(X) * np.cos(Y) * np.cos(beta * Z)
    v = -np.cos(X) * np.sin(Y) * np.cos(beta * Z)
    w = np.zeros_like(u)
    # A_c = sum a cos(kx x + ky y + kz z + phi); curl A analytically
    du = np.zeros_like(u); dv = np.zeros_like(u); dw = np.zeros_like(u)
    for _ in range(12):
        k = rng.integers(-3, 4, size=3)
        if not k.any():
            continue
        a = rng.standard_normal(3)
        phi = rng.uniform(0, 2 * np.pi)
        arg = k[0] * X + k[1] * Y + k[2] * beta * Z + phi
        s = -np.sin(arg)     # d/d(arg) of cos
        kz = k[2] * beta
        # curl A = (dAz/dy - dAy/dz, dAx/dz - dAz/dx, dAy/dx - dAx/dy)
        du += (a[2] * k[1] - a[1] * kz) * s
        dv += (a[0] * kz - a[2] * k[0]) * s
        dw += (a[1] * k[0] - a[0] * k[1]) * s
    scale = amp / max(np.abs(du).max(), np.abs(dv).max(), np.abs(dw).max(), 1e-300)
    fields = []
    for f, d in ((u, du), (v, dv), (w, dw)):
        full = np.zeros((nz, mesh.planeSize))
        full[:, :mesh.nPlane] = fourier_pack(f + scale * d)
        fields.append(full)
    return fields


def batch_problems(count, seed=4):
    """Independent 2D Dirichlet problems: K x K uniform quads on
    [0, 2 pi]^2 with K ~ U{4..16}, np ~ U{5..11}, lambda2 ~ U[0.1, 100]
    (numpy PCG64 stream of `seed`).  Returns [(np, K, lambda2, coeff_seed)]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        K = int(rng.integers(4, 17))
        np_ = int(rng.integers(5, 12))
        lam = float(rng.uniform(0.1, 100.0))
        out.append((np_, K, lam, int(rng.integers(0, 2**31 - 1))))
    return out


def single_plane_problem(mesh, lambda2, seed):
    """f, BC lines and exact solution of lap u - lambda2 u = f for one
    seeded smooth u* on a 2D mesh (nz = 1)."""
    return helmholtz_batch(mesh, 1, lambda2, 0.0, seed)
