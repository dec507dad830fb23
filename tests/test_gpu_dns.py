"""GPU parity of the DNS-path field operators and time step against the CPU
FP64 oracle and the reference's regress golden values, through the C ABI.

Bars (floating point, FP64):
- dDFTr (femlib/fourier.c:47-171): rel L2 <= 1e-13 vs the oracle's FFT, the
  inverse of the forward transform returns the input to 1e-13;
- gradient / smooth (src/auxfield.cpp:514-665, src/field.cpp:190-236):
  rel L2 <= 1e-13 vs the oracle;
- integrate (dns/integrate.cpp:79-221): regress/taylor2.ok and taylor3.ok
  error norms to all printed digits, and rel L2 <= 1e-12 vs the oracle's
  own time step after 10 steps of a perturbed 3D Taylor-Green flow (the
  pressure after removing the constant of its pinned k=0 mode).
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
import semtex_amd as sx
from semtex_amd.mesh import rect_mesh
from solve_helpers import rel

pytestmark = pytest.mark.gpu


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64).ravel().copy()).to("cuda:0")


@pytest.mark.parametrize("N", [2, 6, 8, 12, 30, 64, 96, 128, 160, 256])
def test_dft_r_matches_oracle(cuda, N):
    import torch
    print(f"[dft_r] N={N}", flush=True)
    ntrn = 2 * 1237
    rng = np.random.default_rng(N)
    x = rng.standard_normal((N, ntrn))
    for sign in (1, -1):
        xo = x.copy()
        O.dftr(xo, N, ntrn, sign)
        t = dev(torch, x)
        sx.dft_r(t, N, ntrn, sign)
        torch.cuda.synchronize()
        r = rel(t.cpu().numpy().reshape(N, ntrn), xo)
        print(f"  sign {sign:+d}: rel L2 vs oracle {r:.3e}", flush=True)
        assert r <= 1e-13
    t = dev(torch, x)
    sx.dft_r(t, N, ntrn, 1)
    sx.dft_r(t, N, ntrn, -1)
    torch.cuda.synchronize()
    assert rel(t.cpu().numpy().reshape(N, ntrn), x) <= 1e-13


def test_dft_r_rejects_bad_sizes(cuda):
    import torch
    t = torch.zeros(7 * 4, dtype=torch.float64, device="cuda:0")
    with pytest.raises(sx.SemxError):
        sx.dft_r(t, 7, 4, 1)          # 7 is not 2^p 3^q 5^r
    t = torch.zeros(8 * 3, dtype=torch.float64, device="cuda:0")
    with pytest.raises(sx.SemxError):
        sx.dft_r(t, 8, 3, 1)          # odd number of lines


def mesh_pair(np_, **kw):
    m = rect_mesh(**kw)
    return sx.Mesh(np_, m), O.Mesh(np_, m)


@pytest.mark.parametrize("direction", [0, 1, 2])
def test_gradient_matches_oracle(cuda, direction):
    import torch
    print(f"[gradient] dir {direction}", flush=True)
    P, Q = mesh_pair(7, nx=4, ny=3, distort=0.3, bc=("D", "N", "D", "D"))
    nz, beta = 8, 1.7
    rng = np.random.default_rng(3)
    a = rng.standard_normal((nz, P.planeSize))
    ao = a.copy()
    O.lib().orc_gradient(C.c_void_p(Q.h), nz, beta, O.dp(ao), direction)
    t = dev(torch, a)
    sx.gradient(P, nz, beta, t, direction)
    torch.cuda.synchronize()
    g = t.cpu().numpy().reshape(nz, P.planeSize)
    r = rel(g[:, :P.nPlane], ao[:, :P.nPlane])
    print(f"  rel L2 vs oracle {r:.3e}", flush=True)
    assert r <= 1e-13


def test_smooth_matches_oracle(cuda):
    import torch
    P, Q = mesh_pair(6, nx=5, ny=4, distort=0.25, periodic_x=True)
    nz = 4
    rng = np.random.default_rng(9)
    a = rng.standard_normal((nz, P.planeSize))
    ao = a.copy()
    O.lib().orc_smooth(C.c_void_p(Q.h), nz, O.dp(ao))
    t = dev(torch, a)
    sx.smooth(P, nz, t)
    torch.cuda.synchronize()
    g = t.cpu().numpy().reshape(nz, P.planeSize)
    r = rel(g[:, :P.nPlane], ao[:, :P.nPlane])
    print(f"[smooth] rel L2 vs oracle {r:.3e}", flush=True)
    assert r <= 1e-13


@pytest.mark.parametrize("nz", [1, 8])
def test_regress_taylor_on_gpu(cuda, nz):
    """regress/taylor2.ok (nz=1) and taylor3.ok (nz=8): 2x2 periodic
    elements on [0,2]^2, np=11, nu=0.01, dt=0.02, N_TIME=2, alternating
    skew-symmetric terms, 20 steps; |u-u*|, |v-v*| = 1.127e-05 and
    |p-p*| = 4.224e-01 as printed by the reference."""
    import torch
    print(f"[taylor] nz={nz}", flush=True)
    m = rect_mesh(2, 2, 0, 2, 0, 2, periodic_x=True, periodic_y=True)
    P = sx.Mesh(11, m)
    x, y = P.x, P.y
    nu, dt = 0.01, 0.02
    ps = P.planeSize
    u = np.zeros((nz, ps)); v = np.zeros((nz, ps))
    u[0, :P.nPlane] = -np.cos(np.pi * x) * np.sin(np.pi * y)
    v[0, :P.nPlane] = np.sin(np.pi * x) * np.cos(np.pi * y)
    tu, tv, tw, tp = dev(torch, u), dev(torch, v), dev(torch, np.zeros((nz, ps))), \
        dev(torch, np.zeros((nz, ps)))
    D = sx.DNS(P, nz, np.pi, dt, nu, norder=2, advection=1, tol_rel=1e-12, step_max=500)
    D.run(20, tu, tv, tw, tp)
    torch.cuda.synchronize()
    u = tu.cpu().numpy().reshape(nz, ps); v = tv.cpu().numpy().reshape(nz, ps)
    p = tp.cpu().numpy().reshape(nz, ps)
    T = 0.4
    ue = -np.cos(np.pi * x) * np.sin(np.pi * y) * np.exp(-2 * np.pi ** 2 * nu * T)
    ve = np.sin(np.pi * x) * np.cos(np.pi * y) * np.exp(-2 * np.pi ** 2 * nu * T)
    pe = -0.25 * (np.cos(2 * np.pi * x) + np.cos(2 * np.pi * y)) * np.exp(-4 * np.pi ** 2 * nu * T)
    eu = np.abs(u[0, :P.nPlane] - ue).max()
    ev = np.abs(v[0, :P.nPlane] - ve).max()
    ep = np.abs(p[0, :P.nPlane] - pe).max()
    print(f"  |u-u*| {eu:.3e} |v-v*| {ev:.3e} |p-p*| {ep:.3e}", flush=True)
    assert f"{eu:.3e}" == "1.127e-05" and f"{ev:.3e}" == "1.127e-05"
    assert f"{ep:.3e}" == "4.224e-01"
    assert D.steps == 20


def perturbed_taylor_green(P, nz, beta, seed=7, amp=1e-2):
    """Fourier-space u, v, w of u = sin x cos y cos z, v = -cos x sin y cos z
    on [0,2pi]^3 plus a seeded low-mode perturbation on modes 0..2."""
    rng = np.random.default_rng(seed)
    ps = P.planeSize
    x, y = P.x, P.y
    fields = [np.zeros((nz, ps)) for _ in range(3)]
    # cos(beta z) is mode 1, real part 1/2 in semtex's packing (plane 2)
    fields[0][2, :P.nPlane] = 0.5 * np.sin(x) * np.cos(y)
    fields[1][2, :P.nPlane] = -0.5 * np.cos(x) * np.sin(y)
    for f in fields:
        for k in (0, 2, 3, 4, 5):
            a, b, c, d = rng.standard_normal(4)
            f[k, :P.nPlane] += amp * (a * np.sin(x + c) * np.cos(y + d) + b * np.cos(2 * x + d))
    return fields


@pytest.mark.parametrize("nx", [4, 3], ids=["4x4", "3x3-odd-plane"])
def test_dns_matches_oracle_3d(cuda, nx):
    """10 steps of a perturbed 3D Taylor-Green flow against the oracle, every
    advection form (skew-symmetric, alternating, convective, Stokes).  On 3x3 elements of np=7 a plane holds 441 values and
    is padded to 442, which exercises every kernel's plane stride and the
    padding entries of the fused nonlinear terms."""
    import torch
    m = rect_mesh(nx, nx, 0, 2 * np.pi, 0, 2 * np.pi, periodic_x=True, periodic_y=True)
    np_, nz, beta, dt, nu = 7, 8, 1.0, 2e-3, 0.05
    P = sx.Mesh(np_, m)
    Q = O.Mesh(np_, m)
    u, v, w = perturbed_taylor_green(P, nz, beta)
    p = np.zeros_like(u)
    nsteps = 10
    for adv in (0, 1, 2, 3):
        print(f"[dns 3d] {nx}x{nx} advection {adv} planeSize {P.planeSize} nPlane {P.nPlane}",
              flush=True)
        tu, tv, tw, tp = (dev(torch, a) for a in (u, v, w, p))
        D = sx.DNS(P, nz, beta, dt, nu, norder=3, advection=adv, tol_rel=1e-14, step_max=2000)
        D.run(nsteps, tu, tv, tw, tp)
        torch.cuda.synchronize()
        uo, vo, wo, po = (np.ascontiguousarray(a.copy()) for a in (u, v, w, p))
        Do = O.lib().orc_dns_create(C.c_void_p(Q.h), nz, beta, dt, nu, 3, adv, 1e-14, 2000, 0)
        t = C.c_double()
        O.lib().orc_dns_run(C.c_void_p(Do), nsteps, O.dp(uo), O.dp(vo), O.dp(wo), O.dp(po),
                            C.byref(t))
        O.lib().orc_dns_destroy(C.c_void_p(Do))
        for name, tg, o in (("u", tu, uo), ("v", tv, vo), ("w", tw, wo)):
            g = tg.cpu().numpy().reshape(nz, P.planeSize)
            keep = [k for k in range(nz) if k != 1]
            r = rel(g[keep, :P.nPlane], o[keep, :P.nPlane])
            print(f"  {name}: rel L2 vs oracle {r:.3e}", flush=True)
            assert r <= 1e-12
        # pressure: mode 0 is pinned at a different node of each numbering
        g = tp.cpu().numpy().reshape(nz, P.planeSize)
        g[0, :P.nPlane] -= (g[0, :P.nPlane] - po[0, :P.nPlane]).mean()
        keep = [k for k in range(nz) if k != 1]
        if adv == 3:   # Stokes: the pressure is the splitting error alone
            d = np.abs(g[keep, :P.nPlane] - po[keep, :P.nPlane]).max()
            print(f"  p: max difference vs oracle {d:.3e}", flush=True)
            assert d <= 1e-12 * np.abs(u).max()
            continue
        r = rel(g[keep, :P.nPlane], po[keep, :P.nPlane])
        print(f"  p: rel L2 vs oracle (mode-0 constant removed) {r:.3e}", flush=True)
        assert r <= 1e-12


@pytest.mark.parametrize("case", ["taylor4", "taylor5"])
def test_regress_taylor_z_on_gpu(cuda, case):
    """regress/taylor4.ok and taylor5.ok: 2D Taylor flow in the x-z (y-z)
    plane on 2x2 periodic elements of [0,2]^2, np=11, Nz=8, Lz=2, nu=0.01,
    dt=0.02, N_TIME=2, alternating skew-symmetric terms, 20 steps.  Fields
    are transformed on the GPU (semx_dft_r) both ways; the Nyquist plane is
    zeroed before the inverse transform as Domain::transform does
    (src/domain.cpp:249-259).  Error norms to all printed digits."""
    import torch
    from test_oracle import TAYLOR_Z, norm_text, taylor_z_physical
    print(f"[{case}]", flush=True)
    m = rect_mesh(2, 2, 0, 2, 0, 2, periodic_x=True, periodic_y=True)
    P = sx.Mesh(11, m)
    nz, nu, dt, nstep = 8, 0.01, 0.02, 20
    ps = P.planeSize
    u, v, w, _ = taylor_z_physical(case, P, nz, 0.0, nu)
    tu, tv, tw = dev(torch, u), dev(torch, v), dev(torch, w)
    tp = dev(torch, np.zeros((nz, ps)))
    for t in (tu, tv, tw):
        sx.dft_r(t, nz, ps, 1)
    D = sx.DNS(P, nz, np.pi, dt, nu, norder=2, advection=1, tol_rel=1e-12, step_max=500)
    D.run(nstep, tu, tv, tw, tp)
    exact = taylor_z_physical(case, P, nz, nstep * dt, nu)
    got = []
    for t, e in zip((tu, tv, tw, tp), exact):
        t.view(nz, ps)[1].zero_()
        sx.dft_r(t, nz, ps, -1)
        torch.cuda.synchronize()
        f = t.cpu().numpy().reshape(nz, ps)
        got.append(norm_text(np.abs(f[:, :P.nPlane] - e[:, :P.nPlane]).max()))
    print(f"  norms u v w p: {got}", flush=True)
    assert tuple(got) == TAYLOR_Z[case][1]
    assert D.steps == nstep


@pytest.mark.parametrize("level", [1, 2])
def test_dns_iterative_matches_oracle(cuda, level):
    """The ITERATIVE token (dns/integrate.cpp:317-362, semx_dns_set_iterative):
    velocity by JACPCG, at level 2 the pressure by MIXED, PCG to 1e-14:
    rel L2 <= 1e-10 against the oracle's run at the same level after 8
    steps of the perturbed 3D Taylor-Green flow, eager and with the steady
    steps (PCG launches included) replayed from HIP graphs."""
    import torch
    m = rect_mesh(4, 4, 0, 2 * np.pi, 0, 2 * np.pi, periodic_x=True, periodic_y=True)
    np_, nz, beta, dt, nu, nsteps = 7, 8, 1.0, 2e-3, 0.05, 8
    P, Q = sx.Mesh(np_, m), O.Mesh(np_, m)
    u, v, w = perturbed_taylor_green(P, nz, beta)
    p = np.zeros_like(u)
    uo, vo, wo, po = (np.ascontiguousarray(a.copy()) for a in (u, v, w, p))
    Do = O.lib().orc_dns_create(C.c_void_p(Q.h), nz, beta, dt, nu, 3, 1, 1e-14, 5000, 0)
    assert O.lib().orc_dns_set_iterative(C.c_void_p(Do), level) == 0
    t = C.c_double()
    O.lib().orc_dns_run(C.c_void_p(Do), nsteps, O.dp(uo), O.dp(vo), O.dp(wo), O.dp(po),
                        C.byref(t))
    O.lib().orc_dns_destroy(C.c_void_p(Do))
    keep = [k for k in range(nz) if k != 1]
    for graphs in (False, True):
        print(f"[dns iterative] level {level} graphs {graphs}", flush=True)
        tu, tv, tw, tp = (dev(torch, a) for a in (u, v, w, p))
        D = sx.DNS(P, nz, beta, dt, nu, norder=3, advection=1, tol_rel=1e-14, step_max=5000)
        D.set_iterative(level)
        D.set_graph(graphs)
        D.run(nsteps, tu, tv, tw, tp)
        torch.cuda.synchronize()
        with pytest.raises(sx.SemxError):
            D.set_iterative(0)          # only before the first step
        if graphs:
            ng, rp = D.graph_info()
            print(f"  graphs {ng}, replayed steps {rp}", flush=True)
            assert rp == nsteps - 3
        for name, tg, o in (("u", tu, uo), ("v", tv, vo), ("w", tw, wo), ("p", tp, po)):
            g = tg.cpu().numpy().reshape(nz, P.planeSize)
            if name == "p":   # mode 0 pinned at a different node of each numbering
                g[0, :P.nPlane] -= (g[0, :P.nPlane] - o[0, :P.nPlane]).mean()
            r = rel(g[keep, :P.nPlane], o[keep, :P.nPlane])
            print(f"  {name}: rel L2 vs oracle {r:.3e}", flush=True)
            assert r <= 1e-10
        del D
