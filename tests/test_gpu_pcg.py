"""GPU parity of the Jacobi-PCG Helmholtz solve (Field::solve JACPCG branch,
src/field.cpp:627-720) and of the MIXED method (matrix.cpp:93-96) against
the CPU FP64 oracle, through the C ABI.

Bars:
- same loop: with the step limit reached after a few iterations, the GPU
  and the oracle hold the same iterate up to rounding (rel L2 <= 1e-11);
- converged at TOL_REL = 1e-12: the same iteration count (+-1) and a
  distance to the exact discrete (DIRECT) solution no more than 10x the
  oracle PCG's own (floor 1e-12 relative);
- converged at TOL_REL = 1e-13, FP64, or FP32 inner PCG (element operator
  in FP32, or on the BF16 matrix cores with FP32 accumulation) with FP64
  refinement: rel L2 <= 1e-10 vs the oracle DIRECT solution.
"""
import numpy as np
import pytest

import semtex_amd as sx
from solve_helpers import (PERIODIC, drop_mode0_constant, gpu_solve, make_case,
                           oracle_solve, rel, solved)

pytestmark = pytest.mark.gpu

DDDD = ("D", "D", "D", "D")
CASES = [
    # name, nx, np, nz, lambda2, beta, distort, bc
    ("natural-2d", 4, 8, 1, 2.5, 1.0, 0.0, ("D", "N", "D", "N")),
    ("distorted-3d", 4, 6, 8, 1.0, 1.3, 0.3, ("D", "N", "D", "D")),
    ("uniform-16x16-np9-nz8", 16, 9, 8, 20000.0, 1.0, 0.0, DDDD),
    ("periodic-singular", 8, 7, 4, 0.0, 1.0, 0.0, PERIODIC),
    # 3x3 elements of np=7: 441 values per plane, padded to 442
    ("odd-plane-3d", 3, 7, 8, 2.0, 1.0, 0.2, ("D", "N", "D", "D")),
]
IDS = [c[0] for c in CASES]


def initial_guess(ue, seed=3):
    rng = np.random.default_rng(seed)
    return np.ascontiguousarray(ue + 0.1 * rng.standard_normal(ue.shape))


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_pcg_same_iterates_at_step_limit(cuda, case):
    import torch
    name, nx, np_, nz, lam2, beta, dist, bc = case
    print(f"[pcg step-limit] {name}", flush=True)
    # a singular mode 0 is pinned at the last unknown: with the oracle's
    # numbering (as semtex's session.num) the pinned node is the same
    P, Q, f, bcl, ue = make_case(nx, np_, nz, lam2, beta, dist, bc,
                                 same_numbering=bc == PERIODIC)
    u0 = initial_guess(ue)
    S = sx.ModalMatrixSys(P, lam2, beta, 0, max(nz // 2, 1), nz, sx.JACPCG)
    with pytest.warns(RuntimeWarning, match="step limit"):
        ug, it = gpu_solve(torch, P, S, f, bcl, u0=u0, tol=1e-14, step_max=5)
    uo, ito = oracle_solve(Q, nz, lam2, beta, f, bcl, method=sx.JACPCG, u0=u0, tol=1e-14,
                           step_max=5)
    assert it == 5 and np.all(ito[[k for k in range(nz) if not (nz > 1 and k == 1)]] == -1)
    r = rel(solved(P, ug), solved(P, uo))
    print(f"  rel L2 after 4 iterations: {r:.3e}", flush=True)
    assert r <= 1e-11


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_pcg_converged_parity(cuda, case):
    import torch
    name, nx, np_, nz, lam2, beta, dist, bc = case
    print(f"[pcg converged] {name}", flush=True)
    P, Q, f, bcl, ue = make_case(nx, np_, nz, lam2, beta, dist, bc)
    u0 = initial_guess(ue)
    S = sx.ModalMatrixSys(P, lam2, beta, 0, max(nz // 2, 1), nz, sx.JACPCG)
    ug, it = gpu_solve(torch, P, S, f, bcl, u0=u0, tol=1e-12)
    uo, ito = oracle_solve(Q, nz, lam2, beta, f, bcl, method=sx.JACPCG, u0=u0, tol=1e-12)
    ud, _ = oracle_solve(Q, nz, lam2, beta, f, bcl, method=sx.DIRECT)
    print(f"  iterations gpu {it}, oracle {ito.max()}", flush=True)
    assert abs(it - int(ito.max())) <= 1
    if S.info["singular0"]:
        drop_mode0_constant(P, ug, ud)
        drop_mode0_constant(P, uo, ud)
    eg = rel(solved(P, ug), solved(P, ud))
    eo = rel(solved(P, uo), solved(P, ud))
    print(f"  rel L2 vs DIRECT: gpu {eg:.3e}, oracle PCG {eo:.3e}", flush=True)
    assert eg <= max(10 * eo, 1e-12)


def set_precision(S, prec):
    """Inner tolerance and outer limit per precision: the BF16 operator is
    accurate to about 1e-2, so each refinement step gains about that much."""
    if prec == sx.PCG_BF16_REFINED:
        S.set_pcg(prec, inner_tol=1e-2, max_outer=80)
    else:
        S.set_pcg(prec, inner_tol=1e-4, max_outer=30)


@pytest.mark.parametrize("method,prec", [(sx.JACPCG, sx.PCG_FP64),
                                         (sx.JACPCG, sx.PCG_FP32_REFINED),
                                         (sx.JACPCG, sx.PCG_BF16_REFINED),
                                         (sx.MIXED, sx.PCG_FP64),
                                         (sx.MIXED, sx.PCG_FP32_REFINED),
                                         (sx.MIXED, sx.PCG_BF16_REFINED)],
                         ids=["jacpcg-fp64", "jacpcg-fp32r", "jacpcg-bf16r", "mixed-fp64",
                              "mixed-fp32r", "mixed-bf16r"])
@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_pcg_tight_vs_direct(cuda, case, method, prec):
    import torch
    name, nx, np_, nz, lam2, beta, dist, bc = case
    print(f"[pcg tight] {name} method {method} prec {prec}", flush=True)
    P, Q, f, bcl, ue = make_case(nx, np_, nz, lam2, beta, dist, bc)
    S = sx.ModalMatrixSys(P, lam2, beta, 0, max(nz // 2, 1), nz, method)
    set_precision(S, prec)
    ug, it = gpu_solve(torch, P, S, f, bcl, tol=1e-13)
    ud, _ = oracle_solve(Q, nz, lam2, beta, f, bcl, method=sx.DIRECT)
    if S.info["singular0"]:
        drop_mode0_constant(P, ug, ud)
    e = rel(solved(P, ug), solved(P, ud))
    print(f"  iterations {it}, rel L2 vs DIRECT {e:.3e}", flush=True)
    assert e <= 1e-10


@pytest.mark.parametrize("op", ["block", "wave"])
@pytest.mark.parametrize("prec", [sx.PCG_FP64, sx.PCG_FP32_REFINED], ids=["fp64", "fp32r"])
@pytest.mark.parametrize("case", [CASES[1], CASES[4]], ids=[CASES[1][0], CASES[4][0]])
def test_pcg_element_operators(cuda, case, prec, op, monkeypatch):
    """Both scalar element operators the first solve chooses between
    (block-wide, wavefront-level; forced through SEMX_PCG_OP), converged at
    TOL_REL 1e-13: rel L2 <= 1e-10 vs the oracle DIRECT solution."""
    import torch
    monkeypatch.setenv("SEMX_PCG_OP", op)
    name, nx, np_, nz, lam2, beta, dist, bc = case
    print(f"[pcg operator] {name} prec {prec} {op}", flush=True)
    P, Q, f, bcl, ue = make_case(nx, np_, nz, lam2, beta, dist, bc)
    S = sx.ModalMatrixSys(P, lam2, beta, 0, max(nz // 2, 1), nz, sx.JACPCG)
    set_precision(S, prec)
    ug, it = gpu_solve(torch, P, S, f, bcl, tol=1e-13)
    ud, _ = oracle_solve(Q, nz, lam2, beta, f, bcl, method=sx.DIRECT)
    e = rel(solved(P, ug), solved(P, ud))
    print(f"  iterations {it}, rel L2 vs DIRECT {e:.3e}", flush=True)
    assert e <= 1e-10


@pytest.mark.parametrize("method", [sx.DIRECT, sx.MIXED], ids=["direct", "mixed"])
def test_base_mode_share(cuda, method):
    """A system for modes base..base+nmodes-1 (one process's share in
    semtex, Geometry::baseMode = 2): plane 1 is Im of mode 2 and is solved
    (field.cpp:566 skips plane 1 on the root process only).  rel L2 vs the
    oracle's base-mode solve <= 1e-12 (DIRECT) / 1e-10 (MIXED, all modes
    JACPCG since none is global mode 0)."""
    import torch
    import oracle_lib as O
    from semtex_amd.mesh import rect_mesh
    print(f"[base mode] method {method}", flush=True)
    m = rect_mesh(3, 3, distort=0.2, bc=("D", "N", "D", "D"))
    np_, base, nmodes, beta, lam = 6, 2, 2, 1.3, 0.5
    nz = 2 * nmodes
    P, Q = sx.Mesh(np_, m), O.Mesh(np_, m)
    rng = np.random.default_rng(8)
    f = np.zeros((nz, P.planeSize)); f[:, :P.nPlane] = rng.standard_normal((nz, P.nPlane))
    bc = np.ascontiguousarray(rng.standard_normal((nz, P.nbound * np_)))
    S = sx.ModalMatrixSys(P, lam, beta, base, nmodes, nz, method)
    ug, _ = gpu_solve(torch, P, S, f, bc, tol=1e-13)
    uo = np.zeros_like(f)
    O.field_solve_base(Q, nz, lam, beta, base, sx.DIRECT, np.ascontiguousarray(f.copy()),
                       bc, uo)
    assert np.abs(ug[1, :P.nPlane]).max() > 0
    r = rel(ug[:, :P.nPlane], uo[:, :P.nPlane])
    print(f"  rel L2 vs oracle {r:.3e}", flush=True)
    assert r <= (1e-12 if method == sx.DIRECT else 1e-10)


WG_CASES = [(c, W, prec) for c in (CASES[0], CASES[1], CASES[3]) for W in (2, 3, 4)
            for prec in (sx.PCG_FP64, sx.PCG_FP32_REFINED, sx.PCG_BF16_REFINED)]


@pytest.mark.parametrize("case,W,prec", WG_CASES,
                         ids=[f"{c[0]}-w{W}-{['fp64', 'fp32r', 'bf16r'][p]}"
                              for c, W, p in WG_CASES])
def test_pcg_workgroups_per_plane(cuda, case, W, prec, monkeypatch):
    """Each plane's iteration shared by W workgroups (rows, elements and
    assembly rows split, grid barriers, partial sums added in workgroup
    order), forced through SEMX_PCG_WG; converged at TOL_REL 1e-13: rel L2
    <= 1e-10 vs the oracle DIRECT solution, iterations within 2 (+2%) of
    one workgroup per plane, and (FP64) a step-limited solve reported as
    such."""
    import torch
    name, nx, np_, nz, lam2, beta, dist, bc = case
    print(f"[pcg workgroups] {name} W={W} prec {prec}", flush=True)
    P, Q, f, bcl, ue = make_case(nx, np_, nz, lam2, beta, dist, bc)
    its = {}
    for w in (1, W):
        monkeypatch.setenv("SEMX_PCG_WG", str(w))
        S = sx.ModalMatrixSys(P, lam2, beta, 0, max(nz // 2, 1), nz, sx.JACPCG)
        set_precision(S, prec)
        ug, its[w] = gpu_solve(torch, P, S, f, bcl, tol=1e-13)
        assert S.sweep_info()["pcg_w"] == w
    ud, _ = oracle_solve(Q, nz, lam2, beta, f, bcl, method=sx.DIRECT)
    if S.info["singular0"]:
        drop_mode0_constant(P, ug, ud)
    e = rel(solved(P, ug), solved(P, ud))
    print(f"  iterations {its}, rel L2 vs DIRECT {e:.3e}", flush=True)
    assert e <= 1e-10
    assert abs(its[W] - its[1]) <= 2 + its[1] // 50
    if prec == sx.PCG_FP64:     # the step limit is reported by every workgroup count
        with pytest.warns(RuntimeWarning, match="step limit"):
            gpu_solve(torch, P, S, f, bcl, tol=1e-13, step_max=3)
    sx.helm_check(S)


@pytest.mark.parametrize("W,prec", [(1, sx.PCG_BF16_REFINED), (1, sx.PCG_FP64),
                                    (3, sx.PCG_BF16_REFINED), (3, sx.PCG_FP64)],
                         ids=["w1-bf16r", "w1-fp64", "w3-bf16r", "w3-fp64"])
def test_pcg_repeated_solves_bitwise_equal(cuda, W, prec, monkeypatch):
    """The same JACPCG solve repeated from the same data gives the same bits,
    with one and with several workgroups per plane (the vectors are shared
    through global memory inside a workgroup and across its plane's
    workgroups; every barrier first waits for the wavefront's outstanding
    global stores)."""
    import torch
    name, nx, np_, nz, lam2, beta, dist, bc = CASES[1]
    print(f"[pcg repeat] {name} W={W} prec {prec}", flush=True)
    monkeypatch.setenv("SEMX_PCG_WG", str(W))
    P, Q, f, bcl, ue = make_case(nx, np_, nz, lam2, beta, dist, bc)
    S = sx.ModalMatrixSys(P, lam2, beta, 0, max(nz // 2, 1), nz, sx.JACPCG)
    set_precision(S, prec)
    first, its0 = gpu_solve(torch, P, S, f, bcl, tol=1e-12)
    for r in range(4):
        u, its = gpu_solve(torch, P, S, f, bcl, tol=1e-12)
        assert its == its0
        assert np.array_equal(u, first), r


def test_pcg_balanced_workgroups(cuda, monkeypatch):
    """After a solve that reads back the planes' iteration counts, the next
    solves give the planes that took the most iterations more workgroups
    (24x24 quads np=9, Nz=128, lambda2 = 0.5 + k^2: 127 planes, two
    workgroups each at first, the low modes up to four afterwards).  The
    balanced solve converges to the uniform one's solution (TOL_REL 1e-13:
    rel L2 <= 1e-10), within 2 (+2%) of its iterations, and repeats bit for
    bit."""
    import torch
    monkeypatch.delenv("SEMX_PCG_WG", raising=False)
    nx, np_, nz, lam2, beta = 24, 9, 128, 0.5, 1.0
    print(f"[pcg balanced] {nx}x{nx} np={np_} nz={nz}", flush=True)
    P, Q, f, bcl, ue = make_case(nx, np_, nz, lam2, beta, 0.0, DDDD)
    S = sx.ModalMatrixSys(P, lam2, beta, 0, nz // 2, nz, sx.JACPCG)
    u1, it1 = gpu_solve(torch, P, S, f, bcl, tol=1e-13)
    w1 = S.sweep_info()["pcg_w"]
    u2, it2 = gpu_solve(torch, P, S, f, bcl, tol=1e-13)
    w2 = S.sweep_info()["pcg_w"]
    u3, it3 = gpu_solve(torch, P, S, f, bcl, tol=1e-13)
    e = rel(solved(P, u2), solved(P, u1))
    print(f"  workgroups per plane {w1} -> up to {w2}; iterations {it1}, {it2}; "
          f"rel L2 balanced vs uniform {e:.3e}", flush=True)
    assert w1 == 2 and w2 > w1
    assert e <= 1e-10
    assert abs(it2 - it1) <= 2 + it1 // 50
    assert it3 == it2 and np.array_equal(u3, u2)
