"""Cylindrical coordinates (CYLINDRICAL = 1) on the host: the oracle's
restatement of semtex's cylindrical branches (Element::mapping premultiplies
Q1-Q4 by y, src/element.cpp:1417-1426; HelmholtzKern/Row/Diag use
betak2 / r^2 + lambda2, 0 on the axis, :1464-1618; Element::weight keeps
the unweighted mass, :831-837; Edge areas carry y, src/edge.cpp:75) pinned
by:
- poisson2 (regress/poisson2.ok, "noise-level": |error|_inf < 1e-12 as
  utility/compare.cpp:85,234 prints it), which the same mesh solved as
  Cartesian misses by O(1) -- the pin sees the weighting;
- laplace9's exact harmonic solution in an annulus, 3D (N_Z = 32), and
  the same with the end x = 0 natural (its flux weighted by y along the
  side);
- a solid cylinder with the axis (N_Z = 32), modes split as semtex's per-mode
  numberings split them (axis natural for mode 0, essential for k >= 1);
and the library's host-side condensed element matrices equal the oracle's
(orc_helm_sc) for cylindrical meshes, modes 0 and 3, and for the coupled
fields v~ / w~ (Fourier constants beta k + 1 / beta k - 1, 1 for mode 0;
Field::modeConstant, src/field.cpp:1121-1146).  The coupling itself
(AuxField::couple) is pinned by its closed form (v~ = v + i w) and, with
the constants, by fields whose coupled forms solve the coupled viscous
operators exactly (powers of y), which the scalar constant misses."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
import semtex_amd as sx
from cyl_cases import (axis_lines, fourier_exact, laplace9_exact, laplace9_mesh,
                       poisson2_mesh, poisson2_problem, tube_exact, tube_meshes)
from semtex_amd.mesh import rect_mesh


def oracle_mesh(np_, m, cyl=True):
    Q = O.Mesh(np_, m)
    if cyl:
        Q.set_cylindrical()
    return Q


@pytest.mark.parametrize("method", [sx.DIRECT, sx.JACPCG], ids=["direct", "jacpcg"])
def test_poisson2_golden(method):
    Q = oracle_mesh(11, poisson2_mesh())
    f, bc, ex = poisson2_problem(Q)
    u = np.zeros_like(f)
    O.field_solve(Q, 1, 0.0, 1.0, method, f.copy(), bc, u, tol=1e-14)
    err = np.abs(u[0, :Q.nPlane] - ex[0, :Q.nPlane]).max()
    print(f"poisson2 ({'DIRECT' if method == sx.DIRECT else 'JACPCG'}): |c - (1-4y^2)|_inf "
          f"{err:.3e}")
    assert err < 1e-12                            # regress/poisson2.ok: noise-level
    if method == sx.DIRECT:
        # the same mesh as a Cartesian problem solves another equation
        Qc = oracle_mesh(11, poisson2_mesh(), cyl=False)
        u2 = np.zeros_like(f)
        O.field_solve(Qc, 1, 0.0, 1.0, sx.DIRECT, f.copy(), bc, u2)
        assert np.abs(u2[0, :Q.nPlane] - ex[0, :Q.nPlane]).max() > 1e-2


def test_laplace9_exact():
    nz = 32
    Q = oracle_mesh(11, laplace9_mesh())
    F, bc = fourier_exact(Q, nz, laplace9_exact)
    u = np.zeros_like(F)
    O.field_solve(Q, nz, 0.0, 1.0, sx.DIRECT, np.zeros_like(F), bc, u)
    keep = [k for k in range(nz) if k != 1]
    err = np.abs(u[keep, :Q.nPlane] - F[keep, :Q.nPlane]).max()
    print(f"laplace9: max Fourier-coefficient error {err:.3e} (scale {np.abs(F).max():.2f})")
    assert err < 1e-9


def test_annulus_natural_end_exact():
    """The flux through the end x = 0 enters weighted by y along it."""
    nz = 32
    from cyl_cases import annulus_natural_mesh, laplace9_flux_left, natural_lines
    Q = oracle_mesh(11, annulus_natural_mesh())
    F, bc = fourier_exact(Q, nz, laplace9_exact)
    bc = natural_lines(Q, nz, bc, laplace9_flux_left, 3)
    u = np.zeros_like(F)
    O.field_solve(Q, nz, 0.0, 1.0, sx.DIRECT, np.zeros_like(F), bc, u)
    keep = [k for k in range(nz) if k != 1]
    err = np.abs(u[keep, :Q.nPlane] - F[keep, :Q.nPlane]).max()
    print(f"annulus, natural end: max Fourier-coefficient error {err:.3e}")
    assert err < 1e-9


def tube_solve_oracle(nz, method=sx.DIRECT):
    mN, mD = tube_meshes()
    QN, QD = oracle_mesh(11, mN), oracle_mesh(11, mD)
    F, bc = fourier_exact(QN, nz, tube_exact)
    bc = axis_lines(QN, bc, True)
    u = np.zeros_like(F)
    ps, nl = QN.planeSize, QN.nbound * QN.np
    f = np.zeros_like(F)
    u0, u1 = u[:2].copy(), u[2:].copy()
    O.field_solve_base(QN, 2, 0.0, 1.0, 0, method, f[:2].copy(), bc[:2].copy(), u0, tol=1e-13)
    O.field_solve_base(QD, nz - 2, 0.0, 1.0, 1, method, f[2:].copy(), bc[2:].copy(), u1,
                       tol=1e-13)
    u[:2], u[2:] = u0, u1
    return QN, F, u


@pytest.mark.parametrize("method", [sx.DIRECT, sx.JACPCG], ids=["direct", "jacpcg"])
def test_tube_with_axis_exact(method):
    nz = 32
    Q, F, u = tube_solve_oracle(nz, method)
    keep = [k for k in range(nz) if k != 1]
    err = np.abs(u[keep, :Q.nPlane] - F[keep, :Q.nPlane]).max()
    print(f"tube with axis: max Fourier-coefficient error {err:.3e}")
    assert np.abs(F[0, :Q.nPlane]).max() > 0.5 and np.abs(F[2, :Q.nPlane]).max() > 0.1
    assert err < 1e-9


@pytest.mark.parametrize("mode", [0, 3])
def test_cylindrical_condensation_matches_oracle(mode):
    np_, lam, beta = 7, 2.5, 1.0
    m = rect_mesh(3, 3, 0.0, 2.0, 0.0, 1.0, distort=0.2, bc=("N", "D", "D", "D"))
    P = sx.Mesh(np_, m, cylindrical=True)
    Q = oracle_mesh(np_, m)
    S = sx.ModalMatrixSys(P, lam, beta, 0, 4, 8, sx.DIRECT)
    bk2 = (beta * mode) ** 2
    for e in (0, 4, 8):
        hbb, hbi, hii = S.element(mode, e)
        ob = np.zeros((P.next, P.next))
        oi = np.zeros((P.next, P.nint))
        oii = np.zeros((P.nint, P.nint))
        O.lib().orc_helm_sc(C.c_void_p(Q.h), e, C.c_double(lam), C.c_double(bk2),
                            O.dp(ob), O.dp(oi), O.dp(oii))
        for a, b in ((hbb, ob), (hbi, oi), (hii, oii)):
            assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


def test_cylindrical_mesh_checks():
    m = rect_mesh(2, 2, 0.0, 1.0, -0.5, 0.5)
    with pytest.raises(sx.SemxError):
        sx.Mesh(5, m, cylindrical=True)           # nodes at y < 0
    P = sx.Mesh(5, rect_mesh(2, 2))
    sx.ModalMatrixSys(P, 1.0, 1.0, 0, 1, 1, sx.DIRECT)
    with pytest.raises(sx.SemxError):             # a system already built on it
        sx.check(sx.load().semx_mesh_set_cylindrical(P.h), "semx_mesh_set_cylindrical")


def test_couple_closed_form():
    """AuxField::couple: v~ = v + i w, w~ = v - i w on each mode's (Re, Im)
    planes from mode 1 (mode 0 too when the planes start above it), and the
    inverse undoes it."""
    rng = np.random.default_rng(3)
    nz, n = 8, 13
    v0, w0 = rng.standard_normal((nz, n)), rng.standard_normal((nz, n))
    for base in (0, 1):
        v, w = v0.copy(), w0.copy()
        O.lib().orc_couple(O.dp(v), O.dp(w), nz, n, base, 1)
        for k in range(nz // 2):
            V = v0[2 * k] + 1j * v0[2 * k + 1]
            W = w0[2 * k] + 1j * w0[2 * k + 1]
            if k == 0 and base == 0:
                assert np.array_equal(v[:2], v0[:2]) and np.array_equal(w[:2], w0[:2])
                continue
            assert np.allclose(v[2 * k] + 1j * v[2 * k + 1], V + 1j * W, rtol=0, atol=1e-15)
            assert np.allclose(w[2 * k] + 1j * w[2 * k + 1], V - 1j * W, rtol=0, atol=1e-15)
        O.lib().orc_couple(O.dp(v), O.dp(w), nz, n, base, -1)
        assert np.abs(v - v0).max() < 1e-15 and np.abs(w - w0).max() < 1e-15


def coupled_oracle_solve(Q, nz, F, bc, method=sx.DIRECT):
    """v~, w~ formed (couple), each solved with its Fourier constant
    (modeConstant 'v' / 'w'), then uncoupled: the viscous step of
    dns/integrate.cpp for the cylindrical (v, w) pair, lambda2 = 0, no
    forcing."""
    (Fv, Fw), (bv, bw) = F, bc
    bv, bw = bv.copy(), bw.copy()
    O.lib().orc_couple(O.dp(bv), O.dp(bw), nz, bv.shape[1], 0, 1)
    out = []
    for name, b in (("v", bv), ("w", bw)):
        O.lib().orc_set_field_name(ord(name))
        u = np.zeros_like(Fv)
        O.field_solve(Q, nz, 0.0, 1.0, method, np.zeros_like(Fv), np.ascontiguousarray(b), u)
        out.append(u)
    O.lib().orc_set_field_name(ord("c"))
    uv, uw = out
    O.lib().orc_couple(O.dp(uv), O.dp(uw), nz, uv.shape[1], 0, -1)
    return uv, uw


def test_coupled_fields_modal_exact():
    """Fields whose coupled forms are powers of y (and ln y) mode by mode:
    with their BC values coupled, v~ and w~ solved with beta k + 1 and
    beta k - 1 and uncoupled, they come back to 1e-8 (ln y is not a
    polynomial: 1e-9 at np = 9; swapping the +-1 of v~ gives 0.29);
    solving both with beta k (the scalar constant) does not."""
    from cyl_cases import coupled_modal_exact, swirl_annulus_mesh
    from semtex_amd.mesh import bc_lines
    nz = 6
    Q = oracle_mesh(9, swirl_annulus_mesh())
    Fv, Fw = coupled_modal_exact(Q, nz)
    bv = np.ascontiguousarray(bc_lines(Fv[:, :Q.nPlane], Q.belmt, Q.bside, Q.np))
    bw = np.ascontiguousarray(bc_lines(Fw[:, :Q.nPlane], Q.belmt, Q.bside, Q.np))
    uv, uw = coupled_oracle_solve(Q, nz, (Fv, Fw), (bv, bw))
    err = max(np.abs(uv[:, :Q.nPlane] - Fv[:, :Q.nPlane]).max(),
              np.abs(uw[:, :Q.nPlane] - Fw[:, :Q.nPlane]).max())
    print(f"coupled fields, modal exact solution: error {err:.3e}")
    assert err < 1e-8
    # the scalar constant misses it
    u2 = np.zeros_like(Fv)
    O.field_solve(Q, nz, 0.0, 1.0, sx.DIRECT, np.zeros_like(Fv), bv, u2)
    assert np.abs(u2[2:, :Q.nPlane] - Fv[2:, :Q.nPlane]).max() > 1e-3


@pytest.mark.parametrize("field", ["v", "w"])
def test_field_mode_constant_condensation(field):
    """The coupled fields' systems: betak = beta k +- 1 (1 for mode 0)."""
    np_, lam, beta = 6, 1.5, 2.0
    m = rect_mesh(2, 2, 0.0, 1.0, 0.2, 1.0, distort=0.1)
    P = sx.Mesh(np_, m, cylindrical=True)
    Q = oracle_mesh(np_, m)
    S = sx.ModalMatrixSys(P, lam, beta, 0, 3, 6, sx.DIRECT, field=field)
    for mode in (0, 1, 2):
        bk = 1.0 if mode == 0 else beta * mode + (1.0 if field == "v" else -1.0)
        hbb, hbi, hii = S.element(mode, 1)
        ob = np.zeros((P.next, P.next))
        oi = np.zeros((P.next, P.nint))
        oii = np.zeros((P.nint, P.nint))
        O.lib().orc_helm_sc(C.c_void_p(Q.h), 1, C.c_double(lam), C.c_double(bk * bk),
                            O.dp(ob), O.dp(oi), O.dp(oii))
        for a, b in ((hbb, ob), (hbi, oi), (hii, oii)):
            assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()
