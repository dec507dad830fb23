"""Pins of the CPU FP64 oracle (no GPU).

Each check fixes the oracle by something other than its own formulas: a
closed form, a textbook identity, a dense solve of the assembled system, a
library routine (numpy FFT), or a golden value printed in the reference's
regress/*.ok files.  A plausible mistake (a dropped term, a transposed
operand, a wrong sign or index) fails at least one of them.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from semtex_amd.mesh import bc_lines, rect_mesh, side_nodes


# --------------------------------------------------------------------------
# a2: GLL quadrature and derivative (femlib/polylib.c:312-340, 513-556)
# --------------------------------------------------------------------------

def test_gll_known_answer_np3():
    z, w, D, DT = O.gll(3)
    assert np.allclose(z, [-1, 0, 1], atol=1e-15)
    assert np.allclose(w, [1 / 3, 4 / 3, 1 / 3], atol=1e-15)
    # Lagrange derivative at {-1,0,1}
    assert np.allclose(D, [[-1.5, 2, -0.5], [-0.5, 0, 0.5], [0.5, -2, 1.5]], atol=1e-14)
    assert np.allclose(DT, D.T)


@pytest.mark.parametrize("np_", [3, 5, 8, 9, 11, 13])
def test_gll_exactness(np_):
    N = np_ - 1
    z, w, D, _ = O.gll(np_)
    # points: +-1 and the roots of P_N' (numpy Legendre)
    dP = np.polynomial.legendre.Legendre.basis(N).deriv()
    assert np.allclose(np.sort(z[1:-1]), np.sort(dP.roots().real), atol=1e-13)
    # quadrature exact to degree 2N-1, not to 2N
    for k in range(2 * N):
        exact = (1 - (-1) ** (k + 1)) / (k + 1)
        assert abs(w @ z ** k - exact) < 1e-13
    assert abs(w @ z ** (2 * N) - 2 / (2 * N + 1)) > 1e-9
    # differentiation exact to degree N
    for k in range(N + 1):
        d = D @ z ** k
        ref = k * z ** (k - 1) if k else np.zeros_like(z)
        assert np.allclose(d, ref, atol=1e-11 * max(1, k * k))
    # closed-form corner entries
    assert abs(D[0, 0] + N * (N + 1) / 4) < 1e-12
    assert abs(D[N, N] - N * (N + 1) / 4) < 1e-12


def test_legendre_transform_roundtrip():
    for np_ in (5, 9, 12):
        FW = np.zeros((np_, np_)); BW = np.zeros((np_, np_))
        O.lib().orc_legtran(np_, O.dp(FW), O.dp(BW))
        assert np.allclose(BW @ FW, np.eye(np_), atol=1e-13)
        # BW columns are Legendre polynomials at the GLL points
        z, _, _, _ = O.gll(np_)
        for k in range(np_):
            Pk = np.polynomial.legendre.Legendre.basis(k)(z)
            assert np.allclose(BW[:, k], Pk, atol=1e-13)


def test_svv_operator():
    np_ = 9
    _, _, D, _ = O.gll(np_)
    S, ST = O.svv_operators(np_)             # SVV off: S == 1
    assert np.allclose(S, D, atol=1e-12)
    assert np.allclose(ST, S.T)
    S2, _ = O.svv_operators(np_, mn=4, eps=0.5, nu=0.01)
    z, _, _, _ = O.gll(np_)
    # modes up to SVV_MN are left untouched: exact derivative of z^k, k <= 4
    for k in range(5):
        ref = k * z ** (k - 1) if k else np.zeros_like(z)
        assert np.allclose(S2 @ z ** k, ref, atol=1e-10)
    # the top mode is amplified by sqrt(1 + eps/nu)
    PN = np.polynomial.legendre.Legendre.basis(np_ - 1)(z)
    assert not np.allclose(S2 @ PN, D @ PN, atol=1e-6)


@pytest.mark.parametrize("np_,mn,eps,nu", [(9, 4, 0.5, 0.01), (11, 6, 0.1, 0.002), (7, 0, 1.0, 1.0)])
def test_svv_operator_closed_form(np_, mn, eps, nu):
    """SDV = BW diag(sqrt S) FW DV (src/svv.cpp:126-185) on Legendre
    polynomials: P_m' = sum_{j = m-1, m-3, ... >= 0} (2j+1) P_j, so
    SDV P_m = sum_j (2j+1) sqrt(S_j) P_j at the GLL points, with the filter
    S_j = 1 for j <= mN and 1 + eps/nu exp(-((N-j)/(mN-j))^2) above
    (src/svv.cpp:41-79).  A wrong power of S, a shifted filter index or a
    transposed operator fails this."""
    N = np_ - 1
    z, _, D, _ = O.gll(np_)
    SDV, SDT = O.svv_operators(np_, mn=mn, eps=eps, nu=nu)
    assert np.array_equal(SDT, SDV.T)
    S = np.ones(np_)
    for j in range(mn + 1, np_):
        S[j] = 1.0 + eps / nu * np.exp(-((N - j) / (mn - j)) ** 2)
    L = np.polynomial.legendre.Legendre.basis
    for m in range(np_):
        ref = sum((2 * j + 1) * np.sqrt(S[j]) * L(j)(z) for j in range(m - 1, -1, -2))
        ref = ref if m else np.zeros_like(z)
        assert np.allclose(SDV @ L(m)(z), ref, rtol=1e-12, atol=1e-11 * max(1, np.abs(ref).max()))
    # S_N = 1 + eps/nu exactly: the top mode's derivative coefficient ratio
    assert np.isclose(S[N], 1.0 + eps / nu)


def test_svv_coeffs_z_closed_form():
    """SVV::coeffs_z (src/svv.cpp:82-123): 1 for modes k <= SVV_MZ, exactly
    1 + eps/nu at the last mode k = N (exp(0)), increasing between;
    a process's share (base > 0) takes the same values; all 1 when SVV_MZ
    is negative or not below N, or eps <= EPSDP."""
    nmode, mz, eps, nu = 16, 5, 0.3, 0.01
    S = O.svv_coeffs_z(nmode, nmode, 0, mz, eps, nu)
    assert np.all(S[:mz + 1] == 1.0)
    assert S[nmode - 1] == 1.0 + eps / nu
    assert np.all(np.diff(S[mz:]) >= 0) and np.all(np.diff(S[mz + 1:]) > 0)
    share = O.svv_coeffs_z(4, nmode, 8, mz, eps, nu)
    assert np.array_equal(share, S[8:12])
    for args in ((nmode, nmode, 0, -1, eps, nu), (nmode, nmode, 0, nmode - 1, eps, nu),
                 (nmode, nmode, 0, mz, 0.0, nu)):
        assert np.all(O.svv_coeffs_z(*args) == 1.0)


# --------------------------------------------------------------------------
# a4: edge maps (femlib/mapping.c:46-132) -- brute-force construction
# --------------------------------------------------------------------------

@pytest.mark.parametrize("np_", [3, 6, 9])
def test_edgemaps(np_):
    emap, pmap = O.edgemaps(np_)
    nm = np_ - 1
    ref = []
    for side in range(4):
        ref += side_nodes(np_, side)[:-1]
    ref += [i * np_ + j for i in range(1, nm) for j in range(1, nm)]
    assert list(emap) == ref
    assert np.all(pmap[emap] == np.arange(np_ * np_))


# --------------------------------------------------------------------------
# a5-a8: geometry and elemental operators
# --------------------------------------------------------------------------

def test_mapping_rectangle_closed_form():
    hx, hy, np_ = 0.7, 0.3, 7
    m = rect_mesh(1, 1, 0, hx, 0, hy)
    M = O.Mesh(np_, m)
    f = M.factors()
    _, w, _, _ = O.gll(np_)
    W = np.outer(w, w).ravel()
    J = hx * hy / 4
    assert np.allclose(f["Q1"], W * hy / hx, rtol=1e-13)
    assert np.allclose(f["Q2"], W * hx / hy, rtol=1e-13)
    assert np.allclose(f["Q4"], W * J, rtol=1e-13)
    assert np.all(f["Q3"] == 0) and f["flags"][0, 4] == 0
    assert np.allclose(f["drdx"], 2 / hx) and np.allclose(f["dsdy"], 2 / hy)
    assert f["flags"][0, 1] == 0 and f["flags"][0, 2] == 0   # null cross partials


def dense_element(np_, hx, hy, lam):
    """K_e + lam M_e of a hx x hy rectangle from Kronecker products."""
    _, w, D, _ = O.gll(np_)
    W = np.diag(w)
    A = D.T @ W @ D                   # 1D stiffness
    # node index i*np + j: i = s (y), j = r (x)
    K = (hy / hx) * np.kron(W, A) + (hx / hy) * np.kron(A, W)
    M = (hx * hy / 4) * np.kron(W, W)
    return K + lam * M, M


def test_element_matrix_kronecker():
    np_, hx, hy, lam = 6, 0.5, 0.25, 3.0
    M = O.Mesh(np_, rect_mesh(1, 1, 0, hx, 0, hy))
    H = np.zeros((np_ * np_, np_ * np_))
    O.lib().orc_helm_matrix(C.c_void_p(M.h), 0, C.c_double(lam), C.c_double(0.0), O.dp(H))
    ref, Mm = dense_element(np_, hx, hy, lam)
    assert np.allclose(H, ref, atol=1e-12 * np.abs(ref).max())
    assert np.allclose(H, H.T, atol=1e-13)
    # K 1 = 0: row sums equal lam * M 1
    assert np.allclose(H.sum(1), lam * Mm.sum(1), atol=1e-12)


def test_helm_op_matches_matrix_distorted():
    np_ = 7
    M = O.Mesh(np_, rect_mesh(3, 3, distort=0.35))
    rng = np.random.default_rng(0)
    for e in (0, 4, 8):
        H = np.zeros((np_ ** 2, np_ ** 2))
        O.lib().orc_helm_matrix(C.c_void_p(M.h), e, C.c_double(2.0), C.c_double(1.5), O.dp(H))
        x = rng.standard_normal(np_ ** 2)
        y = np.zeros_like(x)
        O.lib().orc_helm_op(C.c_void_p(M.h), e, C.c_double(2.0), C.c_double(1.5), O.dp(x), O.dp(y))
        assert np.allclose(y, H @ x, atol=1e-11 * np.abs(H).max())
        assert np.allclose(H, H.T, atol=1e-12 * np.abs(H).max())


def test_static_condensation_schur():
    np_ = 6
    M = O.Mesh(np_, rect_mesh(2, 2, distort=0.3))
    e, lam = 3, 4.0
    H = np.zeros((np_ ** 2, np_ ** 2))
    O.lib().orc_helm_matrix(C.c_void_p(M.h), e, C.c_double(lam), C.c_double(0.0), O.dp(H))
    emap, _ = O.edgemaps(np_)
    He = H[np.ix_(emap, emap)]
    nb = M.next
    hbb = np.zeros((nb, nb)); hbi = np.zeros((nb, M.nint)); hii = np.zeros((M.nint, M.nint))
    O.lib().orc_helm_sc(C.c_void_p(M.h), e, C.c_double(lam), C.c_double(0.0),
                        O.dp(hbb), O.dp(hbi), O.dp(hii))
    Bb, Bi, Ii = He[:nb, :nb], He[:nb, nb:], He[nb:, nb:]
    Iinv = np.linalg.inv(Ii)
    assert np.allclose(hii, Iinv, rtol=1e-10, atol=1e-12)
    assert np.allclose(hbi, Bi @ Iinv, atol=1e-11)
    assert np.allclose(hbb, Bb - Bi @ Iinv @ Bi.T, atol=1e-11 * np.abs(Bb).max())


# --------------------------------------------------------------------------
# a9-a14: the solve, against a dense solve of the assembled system
# --------------------------------------------------------------------------

def assemble_dense(M, lam, betak2):
    """Global (all-node) Helmholtz matrix from element matrices, the assembled
    diagonal mass, and each element's global node ids."""
    np_, npnp = M.np, M.np ** 2
    emap, _ = O.edgemaps(np_)
    npts = M.nglobal + M.nel * M.nint
    A = np.zeros((npts, npts)); B = np.zeros((npts, npts))
    f = M.factors()
    gid_all = []
    for e in range(M.nel):
        gid = np.zeros(npnp, dtype=int)
        gid[emap[:M.next]] = M.btog[e * M.next:(e + 1) * M.next]
        gid[emap[M.next:]] = M.nglobal + e * M.nint + np.arange(M.nint)
        H = np.zeros((npnp, npnp))
        O.lib().orc_helm_matrix(C.c_void_p(M.h), e, C.c_double(lam), C.c_double(betak2), O.dp(H))
        A[np.ix_(gid, gid)] += H
        B[gid, gid] += f["Q4"][e * npnp:(e + 1) * npnp]
        gid_all.append(gid)
    return A, B, gid_all


def test_direct_equals_dense_solve_with_essential_and_natural():
    np_, lam = 6, 2.5
    m = rect_mesh(3, 2, 0, 1.5, 0, 1, bc=("D", "N", "D", "N"), distort=0.25)
    M = O.Mesh(np_, m)
    A, B, gids = assemble_dense(M, lam, 0.0)
    npnp = np_ * np_
    rng = np.random.default_rng(7)
    fphys = rng.standard_normal(M.nPlane)
    bc = rng.standard_normal(M.nbound * np_)
    # right-hand side -sum_e M_e f_e + <h, w> on natural sides (element-local
    # mass: f is discontinuous across elements); essential values g
    b = np.zeros(A.shape[0])
    Q4 = M.factors()["Q4"]
    for e, gid in enumerate(gids):
        sl = slice(e * npnp, (e + 1) * npnp)
        np.add.at(b, gid, -Q4[sl] * fphys[sl])
    area = M.barea()
    emap, _ = O.edgemaps(np_)
    ess = {}
    for k, (e, s) in enumerate(zip(M.belmt, M.bside)):
        nodes = side_nodes(np_, s)
        gid = gids[e]
        kind = m["skind"][e, s]
        for i, q in enumerate(nodes):
            if kind == 2:
                b[gid[q]] += bc[k * np_ + i] * area[k * np_ + i]
            else:
                ess[gid[q]] = bc[k * np_ + i]
    fixed = np.array(sorted(ess))
    free = np.setdiff1d(np.arange(A.shape[0]), fixed)
    g = np.array([ess[i] for i in fixed])
    u = np.zeros(A.shape[0]); u[fixed] = g
    u[free] = np.linalg.solve(A[np.ix_(free, free)], b[free] - A[np.ix_(free, fixed)] @ g)
    ref = np.zeros(M.nPlane)
    for e, gid in enumerate(gids):
        ref[e * npnp:(e + 1) * npnp] = u[gid]
    for method in (0, 1):
        ff = np.zeros(M.planeSize); ff[:M.nPlane] = fphys
        uu = np.zeros(M.planeSize)
        O.field_solve(M, 1, lam, 1.0, method, ff, bc.copy(), uu, tol=1e-14)
        tol = 1e-12 if method == 0 else 1e-9
        assert np.abs(uu[:M.nPlane] - ref).max() <= tol * np.abs(ref).max()


def test_pcg_matches_direct_3d():
    m = rect_mesh(3, 3, distort=0.2)
    M = O.Mesh(7, m)
    nz, lam, beta = 6, 1.0, 1.7
    rng = np.random.default_rng(3)
    f = rng.standard_normal((nz, M.planeSize)); f[1] = 0
    bc = rng.standard_normal((nz, M.nbound * 7)); bc[1] = 0
    u1 = np.zeros_like(f); u2 = np.zeros_like(f)
    O.field_solve(M, nz, lam, beta, 0, f.copy(), bc.copy(), u1)
    O.field_solve(M, nz, lam, beta, 1, f.copy(), bc.copy(), u2, tol=1e-13)
    assert np.linalg.norm(u1 - u2) <= 1e-10 * np.linalg.norm(u1)


def test_base_mode_planes_equal_full_field_planes():
    """One process's share of modes (Geometry::baseMode > 0): its planes are
    those of the full field shifted by 2*base, and its plane 1 is Im of mode
    base, which is solved (field.cpp:566 skips plane 1 on the root only).
    Pinned against the base-0 solve of the whole field, for DIRECT and
    MIXED (DIRECT only for global mode 0, matrix.cpp:93-96)."""
    m = rect_mesh(3, 3, distort=0.2)
    M = O.Mesh(6, m)
    base, nmodes, beta, lam = 2, 2, 1.3, 0.5
    nzf = 2 * (base + nmodes)
    rng = np.random.default_rng(8)
    f = rng.standard_normal((nzf, M.planeSize)); f[1] = 0
    bc = rng.standard_normal((nzf, M.nbound * 6)); bc[1] = 0
    for method, tol in ((0, 1e-12), (2, 1e-10)):
        uf = np.zeros_like(f)
        O.field_solve(M, nzf, lam, beta, method, f.copy(), bc.copy(), uf, tol=1e-13)
        sl = slice(2 * base, nzf)
        ub = np.zeros((2 * nmodes, M.planeSize))
        it = O.field_solve_base(M, 2 * nmodes, lam, beta, base, method,
                                np.ascontiguousarray(f[sl]), np.ascontiguousarray(bc[sl]),
                                ub, tol=1e-13)
        assert np.abs(ub[1]).max() > 0
        assert (it > 0).all() if method == 2 else (it == 0).all()
        assert np.linalg.norm(ub - uf[sl]) <= tol * np.linalg.norm(uf[sl])


# --------------------------------------------------------------------------
# Reference regression pins (regress/*.ok, norm_inf vs the exact solution,
# `noise-level` meaning < 1e-12: utility/compare.cpp:85,229-235)
# --------------------------------------------------------------------------

def solve2d(np_, m, lam, f, exact_bc, method=0, tol=1e-12):
    M = O.Mesh(np_, m)
    x, y = M.x, M.y
    ff = np.zeros(M.planeSize); ff[:M.nPlane] = f(x, y)
    u = np.zeros(M.planeSize)
    bc = np.ascontiguousarray(bc_lines(exact_bc(x, y)[None, :], M.belmt, M.bside, np_)[0])
    O.field_solve(M, 1, lam, 1.0, method, ff, bc, u, tol=tol)
    return M, u[:M.nPlane]


def test_regress_helmholtz1():
    ex = lambda x, y: -np.sin(np.pi * x) * np.sin(np.pi * y) / (2 * np.pi ** 2 + 16)
    M, u = solve2d(11, rect_mesh(2, 2), 16.0, lambda x, y: np.sin(np.pi * x) * np.sin(np.pi * y), ex)
    assert np.abs(u - ex(M.x, M.y)).max() < 1e-12


def test_regress_poisson1():
    ex = lambda x, y: -np.sin(np.pi * x) * np.sin(np.pi * y) / (2 * np.pi ** 2)
    M, u = solve2d(11, rect_mesh(2, 2), 0.0, lambda x, y: np.sin(np.pi * x) * np.sin(np.pi * y), ex)
    assert np.abs(u - ex(M.x, M.y)).max() < 1e-12


def test_regress_laplace1():
    ex = lambda x, y: np.sin(2 * np.pi * x) * np.sinh(2 * np.pi * y) / np.sinh(2 * np.pi)
    M, u = solve2d(11, rect_mesh(2, 2), 0.0, lambda x, y: 0 * x, ex)
    assert f"{np.abs(u - ex(M.x, M.y)).max():.3e}" == "1.052e-10"


def test_regress_laplace4_natural():
    ex = lambda x, y: np.sin(np.pi * x) * np.sinh(np.pi * y) / (np.pi * np.cosh(np.pi))
    flux = lambda x, y: np.where(np.abs(y - 1) < 1e-12, np.sin(np.pi * x), ex(x, y))
    M, u = solve2d(12, rect_mesh(2, 2, bc=("D", "D", "N", "D")), 0.0, lambda x, y: 0 * x, flux)
    assert np.abs(u - ex(M.x, M.y)).max() < 1e-12


def test_regress_laplace3_3d():
    """mesh/laplace3: [-1,1]^2 x [0,2pi], N_Z=8, N_P=10, exact
    exp(sqrt(2) x) sin(y) sin(z); regress/laplace3.ok noise-level."""
    M = O.Mesh(10, rect_mesh(2, 2, -1, 1, -1, 1))
    nz = 8
    zz = np.arange(nz) * 2 * np.pi / nz
    ex = np.exp(np.sqrt(2) * M.x)[None, :] * np.sin(M.y)[None, :] * np.sin(zz)[:, None]
    bcf = np.ascontiguousarray(bc_lines(ex, M.belmt, M.bside, 10))
    O.dftr(bcf, nz, bcf.shape[1], 1)
    u = np.zeros((nz, M.planeSize)); f = np.zeros((nz, M.planeSize))
    O.field_solve(M, nz, 0.0, 1.0, 0, f, bcf, u)
    O.dftr(u, nz, M.planeSize, -1)
    assert np.abs(u[:, :M.nPlane] - ex).max() < 1e-12


# --------------------------------------------------------------------------
# a15: dDFTr semantics vs numpy.fft.rfft (femlib/fourier.c:47-171)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("N", [2 * 3, 8, 10, 12, 30, 64, 96, 128])
def test_dftr_vs_numpy(N):
    ntrn = 10
    x = np.random.default_rng(N).standard_normal((N, ntrn))
    d = np.ascontiguousarray(x.copy())
    O.dftr(d, N, ntrn, 1)
    X = np.fft.rfft(x, axis=0) / N
    ref = np.zeros_like(x)
    ref[0] = X[0].real
    ref[1] = X[N // 2].real
    ref[2::2] = X[1:N // 2].real
    ref[3::2] = X[1:N // 2].imag
    assert np.allclose(d, ref, atol=1e-15 * np.sqrt(N))
    O.dftr(d, N, ntrn, -1)
    assert np.allclose(d, x, atol=1e-14)


def test_dftr_rejects_bad_sizes():
    d = np.zeros(14 * 4)
    with pytest.raises(ValueError):
        O.dftr(d, 14, 4, 1)       # factor 7
    d = np.zeros(8 * 3)
    with pytest.raises(ValueError):
        O.dftr(d, 8, 3, 1)        # odd number of transforms


# --------------------------------------------------------------------------
# a16: gradient and smoothing
# --------------------------------------------------------------------------

def test_gradient_exact_for_polynomials():
    np_ = 8
    M = O.Mesh(np_, rect_mesh(3, 2, 0, 1.2, 0, 0.8, distort=0.3))
    x, y = M.x, M.y
    u = np.zeros((1, M.planeSize)); u[0, :M.nPlane] = x ** 3 * y + 2 * y ** 2
    gx = u.copy(); gy = u.copy()
    O.lib().orc_gradient(C.c_void_p(M.h), 1, C.c_double(1.0), O.dp(gx), 0)
    O.lib().orc_gradient(C.c_void_p(M.h), 1, C.c_double(1.0), O.dp(gy), 1)
    # bilinear elements map cubics to degree-6 polynomials: exact at np=8
    assert np.allclose(gx[0, :M.nPlane], 3 * x ** 2 * y, atol=1e-11)
    assert np.allclose(gy[0, :M.nPlane], x ** 3 + 4 * y, atol=1e-11)


def test_gradient_z_fourier():
    M = O.Mesh(5, rect_mesh(2, 2))
    nz, beta = 8, 1.3
    rng = np.random.default_rng(2)
    d = rng.standard_normal((nz, M.planeSize))
    g = d.copy()
    O.lib().orc_gradient(C.c_void_p(M.h), nz, C.c_double(beta), O.dp(g), 2)
    assert np.all(g[:2] == 0)
    for k in range(1, nz // 2):
        assert np.allclose(g[2 * k], -beta * k * d[2 * k + 1])
        assert np.allclose(g[2 * k + 1], beta * k * d[2 * k])


def test_smooth_preserves_continuous_fields_and_is_idempotent():
    np_ = 6
    M = O.Mesh(np_, rect_mesh(3, 3, distort=0.2))
    u = np.zeros((1, M.planeSize)); u[0, :M.nPlane] = np.cos(M.x) * M.y
    v = u.copy()
    O.lib().orc_smooth(C.c_void_p(M.h), 1, O.dp(v))
    assert np.allclose(v, u, atol=1e-14)
    w = np.random.default_rng(1).standard_normal((1, M.planeSize))
    O.lib().orc_smooth(C.c_void_p(M.h), 1, O.dp(w))
    w2 = w.copy()
    O.lib().orc_smooth(C.c_void_p(M.h), 1, O.dp(w2))
    assert np.allclose(w, w2, atol=1e-14)


# --------------------------------------------------------------------------
# a17/a18: the DNS step against regress/taylor2.ok and taylor3.ok
# (u, v norm_inf 1.127e-05, p 4.224e-01 after 20 steps; mesh/taylor2,3)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("nz", [1, 8])
def test_regress_taylor(nz):
    m = rect_mesh(2, 2, 0, 2, 0, 2, periodic_x=True, periodic_y=True)
    M = O.Mesh(11, m)
    x, y = M.x, M.y
    nu, dt = 0.01, 0.02
    ps = M.planeSize
    u = np.zeros((nz, ps)); v = np.zeros((nz, ps)); w = np.zeros((nz, ps)); p = np.zeros((nz, ps))
    u[0, :M.nPlane] = -np.cos(np.pi * x) * np.sin(np.pi * y)
    v[0, :M.nPlane] = np.sin(np.pi * x) * np.cos(np.pi * y)
    D = O.lib().orc_dns_create(C.c_void_p(M.h), nz, np.pi, dt, nu, 2, 1, 1e-12, 500, 0)
    t = C.c_double()
    O.lib().orc_dns_run(C.c_void_p(D), 20, O.dp(u), O.dp(v), O.dp(w) if nz > 1 else None,
                        O.dp(p), C.byref(t))
    O.lib().orc_dns_destroy(C.c_void_p(D))
    T = 0.4
    ue = -np.cos(np.pi * x) * np.sin(np.pi * y) * np.exp(-2 * np.pi ** 2 * nu * T)
    ve = np.sin(np.pi * x) * np.cos(np.pi * y) * np.exp(-2 * np.pi ** 2 * nu * T)
    pe = -0.25 * (np.cos(2 * np.pi * x) + np.cos(2 * np.pi * y)) * np.exp(-4 * np.pi ** 2 * nu * T)
    assert f"{np.abs(u[0, :M.nPlane] - ue).max():.3e}" == "1.127e-05"
    assert f"{np.abs(v[0, :M.nPlane] - ve).max():.3e}" == "1.127e-05"
    assert f"{np.abs(p[0, :M.nPlane] - pe).max():.3e}" == "4.224e-01"
    if nz > 1:
        assert np.abs(w).max() < 1e-12 and np.abs(u[2:]).max() < 1e-12


@pytest.mark.parametrize("adv", [2, 3], ids=["convective", "stokes"])
def test_taylor_other_advection_forms(adv):
    """dns -C (convective, dns/nonlinear.cpp:414-525) and dns -N (Stokes,
    :528-570) on taylor2's flow, pinned by what the exact solution fixes:
    its u . grad u is the gradient of -(cos 2 pi x + cos 2 pi y)/4 e^(-4 pi^2
    nu t), so the velocity decays as e^(-2 pi^2 nu t) in both forms (to the
    time-stepping error of the alternating run, 1.1e-5); the convective
    form's pressure is that potential (up to a constant), and without
    nonlinear terms the pressure stays zero.  A sign error or a stray 1/2 in
    the convective product moves the pressure by O(0.1)."""
    m = rect_mesh(2, 2, 0, 2, 0, 2, periodic_x=True, periodic_y=True)
    M = O.Mesh(11, m)
    x, y = M.x, M.y
    nu, dt, ps = 0.01, 0.02, M.planeSize
    u = np.zeros((1, ps)); v = np.zeros((1, ps)); p = np.zeros((1, ps))
    u[0, :M.nPlane] = -np.cos(np.pi * x) * np.sin(np.pi * y)
    v[0, :M.nPlane] = np.sin(np.pi * x) * np.cos(np.pi * y)
    D = O.lib().orc_dns_create(C.c_void_p(M.h), 1, np.pi, dt, nu, 2, adv, 1e-12, 500, 0)
    t = C.c_double()
    O.lib().orc_dns_run(C.c_void_p(D), 20, O.dp(u), O.dp(v), None, O.dp(p), C.byref(t))
    O.lib().orc_dns_destroy(C.c_void_p(D))
    T = 0.4
    ue = -np.cos(np.pi * x) * np.sin(np.pi * y) * np.exp(-2 * np.pi ** 2 * nu * T)
    pe = -0.25 * (np.cos(2 * np.pi * x) + np.cos(2 * np.pi * y)) * np.exp(-4 * np.pi ** 2 * nu * T)
    eu = np.abs(u[0, :M.nPlane] - ue).max()
    pc = p[0, :M.nPlane] - p[0, :M.nPlane].mean()
    print(f"advection {adv}: |u - u*| {eu:.3e}, |p| {np.abs(pc).max():.3e}, "
          f"|p - p*| (constant removed) {np.abs(pc - (pe - pe.mean())).max():.3e}")
    assert eu < 2e-5
    if adv == 2:
        assert np.abs(pc - (pe - pe.mean())).max() < 1e-4
    else:
        assert np.abs(pc).max() < 1e-7


# --------------------------------------------------------------------------
# a17/a18 with z-variation: regress/taylor4.ok and taylor5.ok (2D Taylor
# flow in the x-z and y-z planes, mesh/taylor4, mesh/taylor5): these are the
# pins of every z-coupled term of the nonlinear products, setPForce and the
# projection.  The reference's comparison (utility/compare.cpp:212-235)
# reads the physical-space fields that Domain::transform writes, which zeroes
# the Nyquist plane before the inverse transform (src/domain.cpp:249-259).
# --------------------------------------------------------------------------

TAYLOR_Z = {
    # case: (exact u, v, w, p as functions of (a, z, t) with a = x or y;
    #        which in-plane coordinate; golden norms u v w p)
    "taylor4": ("x", ("1.099e-05", "noise-level", "1.176e-05", "2.089e-01")),
    "taylor5": ("y", ("noise-level", "1.099e-05", "1.176e-05", "2.135e-01")),
}


def taylor_z_exact(case, x, y, z, t, nu):
    a = x if TAYLOR_Z[case][0] == "x" else y
    f1 = np.exp(-2 * np.pi ** 2 * nu * t)
    f2 = np.exp(-4 * np.pi ** 2 * nu * t)
    s = -np.cos(np.pi * a) * np.sin(np.pi * z) * f1
    w = np.sin(np.pi * a) * np.cos(np.pi * z) * f1
    p = -0.25 * (np.cos(2 * np.pi * a) + np.cos(2 * np.pi * z)) * f2
    zero = np.zeros_like(s)
    return (s, zero, w, p) if TAYLOR_Z[case][0] == "x" else (zero, s, w, p)


def taylor_z_physical(case, M, nz, t, nu):
    """Physical planes (nz, planeSize) of u, v, w, p at time t; z_k = k Lz/nz
    with Lz = 2 (BETA = 2 pi / Lz = pi)."""
    out = [np.zeros((nz, M.planeSize)) for _ in range(4)]
    for k in range(nz):
        z = 2.0 * k / nz
        for f, e in zip(out, taylor_z_exact(case, M.x, M.y, z, t, nu)):
            f[k, :M.nPlane] = e
    return out


def norm_text(e):
    """utility/compare.cpp:232-235 with -n: below 1e-12 prints noise-level."""
    return "noise-level" if e < 1e-12 else f"{e:.3e}"


@pytest.mark.parametrize("case", ["taylor4", "taylor5"])
def test_regress_taylor_z(case):
    m = rect_mesh(2, 2, 0, 2, 0, 2, periodic_x=True, periodic_y=True)
    M = O.Mesh(11, m)
    nz, nu, dt, nstep = 8, 0.01, 0.02, 20
    ps = M.planeSize
    u, v, w, _ = taylor_z_physical(case, M, nz, 0.0, nu)
    for f in (u, v, w):
        O.dftr(f, nz, ps, 1)
    p = np.zeros((nz, ps))
    D = O.lib().orc_dns_create(C.c_void_p(M.h), nz, np.pi, dt, nu, 2, 1, 1e-12, 500, 0)
    t = C.c_double()
    O.lib().orc_dns_run(C.c_void_p(D), nstep, O.dp(u), O.dp(v), O.dp(w), O.dp(p),
                        C.byref(t))
    O.lib().orc_dns_destroy(C.c_void_p(D))
    exact = taylor_z_physical(case, M, nz, nstep * dt, nu)
    got = []
    for f, e in zip((u, v, w, p), exact):
        f[1] = 0.0                                  # Domain::transform
        O.dftr(f, nz, ps, -1)
        got.append(norm_text(np.abs(f[:, :M.nPlane] - e[:, :M.nPlane]).max()))
    assert tuple(got) == TAYLOR_Z[case][1]


# --------------------------------------------------------------------------
# f3: computed high-order pressure BCs (BCmgr::maintainFourier and
# evaluateCNBCp, src/bcmgr.cpp:1039-1337) against regress/kovas1.ok, and
# enumerate's numbering (utility/enumerate.cpp:680-806), which fixes the
# node at which kovas1's singular pressure is pinned: every printed digit
# of the three norms is reproduced only with the -O3 numbering of the
# regression script (test/testregress) and the exact-solution start.
# --------------------------------------------------------------------------

def run_kovas(Mu, Mp, nstep=None):
    from bounded_cases import KOVAS, kovas_fields
    (u, v, p), (bcu, bcv) = kovas_fields(Mu)
    nu = 1.0 / KOVAS["Re"]
    D = O.lib().orc_dns_create_bc(C.c_void_p(Mu.h), C.c_void_p(Mp.h), 1, 1.0, KOVAS["dt"], nu,
                                  KOVAS["norder"], KOVAS["advection"], 1e-12, 500, 0, O.dp(bcu),
                                  O.dp(bcv), None)
    t = C.c_double()
    O.lib().orc_dns_run(C.c_void_p(D), nstep or KOVAS["nstep"], O.dp(u), O.dp(v), None,
                        O.dp(p), C.byref(t))
    O.lib().orc_dns_destroy(C.c_void_p(D))
    return u, v, p


def test_regress_kovas1():
    from bounded_cases import KOVAS, KOVAS_OK, error_norms, kovas_meshes
    mu, mp = kovas_meshes()
    Mu, Mp = O.Mesh(KOVAS["np"], mu), O.Mesh(KOVAS["np"], mp)
    Mu.enumerate(3)
    Mp.enumerate(3)
    assert error_norms(Mu, *run_kovas(Mu, Mp)) == KOVAS_OK
    # a plausible slip -- the pressure pinned at another node (the -O1
    # numbering) -- moves the printed pressure norm
    Mp1 = O.Mesh(KOVAS["np"], mp)
    Mp1.enumerate(1)
    assert error_norms(Mu, *run_kovas(Mu, Mp1))[2] != KOVAS_OK[2]


def bandwidth(Q):
    bw = 0
    for e in range(Q.nel):
        g = [Q.btog[e * Q.next + i] for i in range(Q.next) if not Q.bmask[e * Q.next + i]]
        if g:
            bw = max(bw, max(g) - min(g))
    return bw + 1


@pytest.mark.parametrize("opt", [0, 1, 2, 3])
def test_enumerate_numbering(opt):
    """Each level gives a bijection of the global nodes with the unknowns
    first (sortGid keeps essential nodes last, in traverse order); level 0
    is the traverse order itself; RCM levels never widen the band of the
    traverse order on this mesh, and level 3 (least bandwidth over every
    RCM root) is no wider than levels 1 and 2."""
    from bounded_cases import sheared_meshes
    m = sheared_meshes(5, 4)[0]
    bws = {}
    for lev in (0, 1, 2, 3):
        Q = O.Mesh(5, m)
        Q.enumerate(lev)
        ids = np.unique(Q.btog)
        assert np.array_equal(ids, np.arange(Q.nglobal))
        unk = Q.btog[Q.bmask == 0]
        ess = Q.btog[Q.bmask == 1]
        assert unk.max() < Q.nsolve <= ess.min()
        assert Q.nband == bandwidth(Q)
        bws[lev] = Q.nband
    assert bws[3] <= min(bws[1], bws[2]) and max(bws[1], bws[2], bws[3]) <= bws[0]
    # the traverse order (level 0) agrees with the library's provisional
    # numbering up to the unknowns-first shuffle: same node partition
    Q0, Ql = O.Mesh(5, m), O.Mesh(5, m)
    Q0.enumerate(0)
    pairs = set(zip(Q0.btog.tolist(), Ql.btog.tolist()))
    assert len(pairs) == Q0.nglobal


def box_run_oracle(update):
    """BOX on the oracle, the BC lines re-set before each step (update) or
    left at their t = 0 values; returns (mesh, u, v)."""
    from bounded_cases import BOX, box_lines, box_meshes
    mu, mp = box_meshes()
    Q, Qp = O.Mesh(BOX["np"], mu), O.Mesh(BOX["np"], mp)
    (u, v), (bu, bv) = box_lines(Q, 0.0, BOX["nu"])
    p = np.zeros_like(u)
    D = O.lib().orc_dns_create_bc(C.c_void_p(Q.h), C.c_void_p(Qp.h), 1, 1.0, BOX["dt"],
                                  BOX["nu"], BOX["norder"], 1, 1e-12, 500, 0, O.dp(bu),
                                  O.dp(bv), None)
    t = C.c_double()
    for n in range(BOX["nstep"]):
        if update:
            _, (bu, bv) = box_lines(Q, (n + 1) * BOX["dt"], BOX["nu"])
            O.lib().orc_dns_set_bc(C.c_void_p(D), O.dp(bu), O.dp(bv), None)
        O.lib().orc_dns_run(C.c_void_p(D), 1, O.dp(u), O.dp(v), None, O.dp(p), C.byref(t))
    O.lib().orc_dns_destroy(C.c_void_p(D))
    return Q, u, v


def test_time_dependent_velocity_bcs():
    """Velocity BCs re-evaluated each step (dns/integrate.cpp:196-201): the
    decaying Taylor vortex in a box, its walls moving with the decay, ends
    near the exact solution; with the t = 0 wall values kept it does not
    (the walls then drive a different flow)."""
    from bounded_cases import BOX, box_exact
    T = BOX["nstep"] * BOX["dt"]
    errs = []
    for update in (True, False):
        Q, u, v = box_run_oracle(update)
        ue, ve = box_exact(Q.x, Q.y, T, BOX["nu"])
        errs.append(max(np.abs(u[0, :Q.nPlane] - ue).max(), np.abs(v[0, :Q.nPlane] - ve).max()))
    print(f"time-dependent BCs: error {errs[0]:.3e}; frozen BCs: {errs[1]:.3e}")
    assert errs[0] < 1e-4 and errs[1] > 100 * errs[0]


def test_dns_iterative_equals_direct():
    """The ITERATIVE token (dns/integrate.cpp:317-362): velocity by JACPCG
    (level 1), pressure by MIXED as well (level 2), each PCG solve to
    |r| <= 1e-14 |b|, give the DIRECT run's fields to 1e-10 after 6 steps
    of a 3D flow; the level is refused once a step has run."""
    m = rect_mesh(3, 3, 0, 2 * np.pi, 0, 2 * np.pi, periodic_x=True, periodic_y=True)
    np_, nz, beta, dt, nu = 6, 8, 1.0, 2e-3, 0.05
    Q = O.Mesh(np_, m)
    rng = np.random.default_rng(3)
    base = []
    for c in range(3):
        f = np.zeros((nz, Q.planeSize))
        for k in (0, 2, 3, 4, 5):
            a, b, s, d = rng.standard_normal(4)
            f[k, :Q.nPlane] = 0.3 * (a * np.sin(Q.x + s) * np.cos(Q.y + d) + b * np.cos(2 * Q.x + d))
        base.append(f)
    out = []
    for level in (0, 1, 2):
        u, v, w = (a.copy() for a in base)
        p = np.zeros_like(u)
        D = O.lib().orc_dns_create(C.c_void_p(Q.h), nz, beta, dt, nu, 3, 1, 1e-14, 5000, 0)
        assert O.lib().orc_dns_set_iterative(C.c_void_p(D), level) == 0
        t = C.c_double()
        O.lib().orc_dns_run(C.c_void_p(D), 6, O.dp(u), O.dp(v), O.dp(w), O.dp(p), C.byref(t))
        assert O.lib().orc_dns_set_iterative(C.c_void_p(D), 0) == 1
        O.lib().orc_dns_destroy(C.c_void_p(D))
        out.append((u, v, w, p))
    keep = [k for k in range(nz) if k != 1]
    for level in (1, 2):
        # the pressure's mode 0 is DIRECT (one pinned node) at every level
        for name, a, b in zip("uvwp", out[level], out[0]):
            r = np.linalg.norm(a[keep, :Q.nPlane] - b[keep, :Q.nPlane]) / \
                np.linalg.norm(b[keep, :Q.nPlane])
            print(f"level {level} {name}: rel L2 vs DIRECT {r:.3e}")
            assert r <= 1e-10
    assert np.linalg.norm(out[0][3][keep]) > 1e-6
