"""Independent 2D problems solved together (one workgroup each) against the
CPU FP64 oracle, through the C ABI.  Each problem's JACPCG system is the
one semx_helm_solve uses; the batch kernel builds its right-hand side in
the same workgroup.  Bars: converged at TOL_REL 1e-13 (FP64, or FP32 inner
PCG -- element operator in FP32 or on the BF16 matrix cores -- with FP64
refinement): rel L2 <= 1e-10 vs the oracle DIRECT solution;
same iteration count (+-1) as the single-system JACPCG path."""
import numpy as np
import pytest

import semtex_amd as sx
from semtex_amd.mesh import rect_mesh
from semtex_amd.workloads import batch_problems, single_plane_problem
from solve_helpers import gpu_solve, oracle_solve, rel

import oracle_lib as O

pytestmark = pytest.mark.gpu


def problems(count, seed, bcs=("D", "D", "D", "D"), distort=0.0):
    out = []
    for (np_, K, lam, cs) in batch_problems(count, seed):
        K = min(K, 8)
        m = rect_mesh(K, K, 0, 2 * np.pi, 0, 2 * np.pi, bc=bcs, distort=distort)
        P, Q = sx.Mesh(np_, m), O.Mesh(np_, m)
        f, bc, ue = single_plane_problem(P, lam, cs)
        out.append((P, Q, lam, f, bc, ue))
    return out


BATCH_CASES = [(sx.PCG_FP64, None), (sx.PCG_FP64, "block"), (sx.PCG_FP64, "wave"),
               (sx.PCG_FP32_REFINED, None), (sx.PCG_FP32_REFINED, "block"),
               (sx.PCG_FP32_REFINED, "wave"),
               (sx.PCG_BF16_REFINED, None), (sx.PCG_BF16_REFINED, "wave"),
               (sx.PCG_BF16_REFINED, "mfma")]


@pytest.mark.parametrize("prec,op", BATCH_CASES,
                         ids=[f"{['fp64', 'fp32r', 'bf16r'][p]}-{o or 'tuned'}"
                              for p, o in BATCH_CASES])
def test_batch_matches_oracle(cuda, prec, op, monkeypatch):
    """Every element operator the batch tuner may pick (forced through
    SEMX_BATCH_OP; the BF16 batch offers the wavefront-level and matrix-core
    ones), and the tuner's own choice, which is one operator per (np,
    element count) class."""
    import torch
    if op is None:
        monkeypatch.delenv("SEMX_BATCH_OP", raising=False)
    else:
        monkeypatch.setenv("SEMX_BATCH_OP", op)
    print(f"[batch] prec {prec} operator {op or 'tuned'}", flush=True)
    probs = problems(7, seed=4) + problems(3, seed=9, bcs=("D", "N", "D", "N"), distort=0.2)
    systems = []
    for (P, Q, lam, f, bc, ue) in probs:
        S = sx.ModalMatrixSys(P, lam, 1.0, 0, 1, 1, sx.JACPCG)
        if prec == sx.PCG_BF16_REFINED:
            S.set_pcg(prec, inner_tol=1e-2, max_outer=80)
        else:
            S.set_pcg(prec, inner_tol=1e-4, max_outer=30)
        systems.append(S)
    B = sx.Batch(systems)
    dev = torch.device("cuda:0")
    fs = [torch.from_numpy(p[3].ravel().copy()).to(dev) for p in probs]
    bcs = [torch.from_numpy(p[4].ravel().copy()).to(dev) for p in probs]
    us = [torch.zeros(p[3].size, dtype=torch.float64, device=dev) for p in probs]
    it = B.solve(fs, bcs, us, tol_rel=1e-13)
    torch.cuda.synchronize()
    worst = 0.0
    for (P, Q, lam, f, bc, ue), u, S in zip(probs, us, systems):
        ud, _ = oracle_solve(Q, 1, lam, 1.0, f, bc, method=sx.DIRECT)
        g = u.cpu().numpy().reshape(f.shape)
        r = rel(g[:, :P.nPlane], ud[:, :P.nPlane])
        worst = max(worst, r)
        assert r <= 1e-10
    ops, tms = B.info()
    print(f"  {len(probs)} problems, max iterations {it}, worst rel L2 vs DIRECT {worst:.3e}; "
          f"operators {ops}, tuning {tms:.1f} ms", flush=True)
    if op is not None:
        assert ops == [op] * len(probs)
    else:
        allowed = {"wave", "mfma"} if prec == sx.PCG_BF16_REFINED else {"block", "wave"}
        assert set(ops) <= allowed
        cls = {}
        for (P, *_), o in zip(probs, ops):
            assert cls.setdefault((P.np, P.nel), o) == o


def test_batch_iterations_match_single_solves(cuda):
    import torch
    probs = problems(4, seed=11)
    dev = torch.device("cuda:0")
    systems = [sx.ModalMatrixSys(p[0], p[2], 1.0, 0, 1, 1, sx.JACPCG) for p in probs]
    single = []
    for S, p in zip(systems, probs):
        _, it = gpu_solve(torch, p[0], S, p[3], p[4], tol=1e-12)
        single.append(it)
    B = sx.Batch(systems)
    fs = [torch.from_numpy(p[3].ravel().copy()).to(dev) for p in probs]
    bcs = [torch.from_numpy(p[4].ravel().copy()).to(dev) for p in probs]
    us = [torch.zeros(p[3].size, dtype=torch.float64, device=dev) for p in probs]
    it = B.solve(fs, bcs, us, tol_rel=1e-12)
    print(f"[batch] iterations batch {it}, single {single}", flush=True)
    assert abs(it - max(single)) <= 1


@pytest.mark.parametrize("prec", [sx.PCG_FP32_REFINED, sx.PCG_BF16_REFINED],
                         ids=["fp32r", "bf16r"])
def test_full_batch_sampled_against_oracle(cuda, prec):
    """The measured batch itself: 512 problems, K x K quads with K ~
    U{4..16}, np ~ U{5..11}, lambda2 ~ U[0.1,100], seed 4, Dirichlet, in one
    launch (element operator by the tuner); every 16th problem (32 of them,
    every np and a spread of K) against the oracle's DIRECT solution,
    rel L2 <= 1e-10 at TOL_REL 1e-13."""
    import torch
    from semtex_amd.workloads import batch_problems
    print(f"[batch 512] prec {prec}", flush=True)
    dev = torch.device("cuda:0")
    spec = batch_problems(512, seed=4)
    systems, fs, bcs, keep = [], [], [], []
    for i, (np_, K, lam, cs) in enumerate(spec):
        m = rect_mesh(K, K, 0, 2 * np.pi, 0, 2 * np.pi)
        P = sx.Mesh(np_, m)
        f, bc, _ = single_plane_problem(P, lam, cs)
        S = sx.ModalMatrixSys(P, lam, 1.0, 0, 1, 1, sx.JACPCG)
        if prec == sx.PCG_BF16_REFINED:
            S.set_pcg(prec, inner_tol=1e-2, max_outer=120)
        else:
            S.set_pcg(prec, inner_tol=1e-4, max_outer=40)
        systems.append(S)
        fs.append(torch.from_numpy(f.ravel().copy()).to(dev))
        bcs.append(torch.from_numpy(bc.ravel().copy()).to(dev))
        if i % 16 == 0:
            keep.append((i, P, m, np_, lam, f, bc))
    us = [torch.zeros_like(f) for f in fs]
    B = sx.Batch(systems)
    it = B.solve(fs, bcs, us, tol_rel=1e-13)
    torch.cuda.synchronize()
    assert all(bool(torch.isfinite(u).all()) for u in us)
    worst, nps, ks = 0.0, set(), set()
    for (i, P, m, np_, lam, f, bc) in keep:
        Q = O.Mesh(np_, m)
        ud, _ = oracle_solve(Q, 1, lam, 1.0, f, bc, method=sx.DIRECT)
        g = us[i].cpu().numpy().reshape(f.shape)
        r = rel(g[:, :P.nPlane], ud[:, :P.nPlane])
        worst = max(worst, r)
        nps.add(np_)
        ks.add(int(round(P.nel ** 0.5)))
        assert r <= 1e-10, (i, np_, P.nel, r)
    print(f"  max iterations {it}; {len(keep)} sampled problems (np {sorted(nps)}, "
          f"K {sorted(ks)}), worst rel L2 vs DIRECT {worst:.3e}", flush=True)
    assert len(nps) >= 5 and max(ks) >= 12


@pytest.mark.parametrize("prec,op", [(sx.PCG_BF16_REFINED, "mfma"), (sx.PCG_FP64, None)],
                         ids=["bf16r-mfma", "fp64"])
def test_repeated_solves_bitwise_equal(cuda, prec, op, monkeypatch):
    """The same batch solved again from the same data gives the same bits.
    Regression test for the workgroup barriers of the PCG kernels: without
    a wait for each wavefront's outstanding global stores before them, a
    wavefront could read an element result or vector entry of another one
    before it landed, and a batch of 160 small problems (np 5 and 6, K 4..7)
    differed in several problems on every repeat (now and then a NaN)."""
    import torch
    if op is None:
        monkeypatch.delenv("SEMX_BATCH_OP", raising=False)
    else:
        monkeypatch.setenv("SEMX_BATCH_OP", op)
    print(f"[batch repeat] prec {prec} op {op}", flush=True)
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(11)
    meshes, systems, fs, bcs = {}, [], [], []
    for _ in range(160):
        np_, K = int(rng.integers(5, 7)), int(rng.integers(4, 8))
        lam, cs = float(rng.uniform(0.1, 100.0)), int(rng.integers(0, 2**31 - 1))
        if (np_, K) not in meshes:
            meshes[(np_, K)] = sx.Mesh(np_, rect_mesh(K, K, 0, 2 * np.pi, 0, 2 * np.pi))
        P = meshes[(np_, K)]
        f, bc, _ = single_plane_problem(P, lam, cs)
        S = sx.ModalMatrixSys(P, lam, 1.0, 0, 1, 1, sx.JACPCG)
        if prec == sx.PCG_BF16_REFINED:
            S.set_pcg(prec, inner_tol=1e-2, max_outer=80)
        systems.append(S)
        fs.append(torch.from_numpy(f.ravel().copy()).to(dev))
        bcs.append(torch.from_numpy(bc.ravel().copy()).to(dev))
    B = sx.Batch(systems)
    first = None
    for r in range(6):
        us = [torch.zeros_like(f) for f in fs]
        B.solve(fs, bcs, us, tol_rel=1e-10, step_max=500)
        torch.cuda.synchronize()
        assert all(bool(torch.isfinite(u).all()) for u in us), r
        if first is None:
            first = us
            continue
        diff = [i for i, (a, b) in enumerate(zip(first, us)) if not torch.equal(a, b)]
        print(f"  repeat {r}: problems differing from the first solve {diff[:8]}", flush=True)
        assert not diff


def test_layout_from_measured_iterations(cuda, monkeypatch):
    """After a solve that reads back the iteration counts, the launch layout
    is rebuilt from the measured costs (unknowns x iterations): the next
    solves of a batch with a few large problems (several workgroups each)
    among small ones converge to the first solve's solutions (TOL_REL
    1e-12: rel L2 <= 1e-9 per problem) with iteration counts within 2
    (+2%)."""
    import torch
    monkeypatch.delenv("SEMX_BATCH_WG", raising=False)
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(5)
    meshes, systems, fs, bcs = {}, [], [], []
    for i in range(48):
        np_, K = (9, 16) if i % 12 == 0 else (int(rng.integers(5, 9)), int(rng.integers(4, 10)))
        lam, cs = float(rng.uniform(0.1, 100.0)), int(rng.integers(0, 2**31 - 1))
        if (np_, K) not in meshes:
            meshes[(np_, K)] = sx.Mesh(np_, rect_mesh(K, K, 0, 2 * np.pi, 0, 2 * np.pi))
        P = meshes[(np_, K)]
        f, bc, _ = single_plane_problem(P, lam, cs)
        systems.append(sx.ModalMatrixSys(P, lam, 1.0, 0, 1, 1, sx.JACPCG))
        fs.append(torch.from_numpy(f.ravel().copy()).to(dev))
        bcs.append(torch.from_numpy(bc.ravel().copy()).to(dev))
    B = sx.Batch(systems)
    sols, its = [], []
    for r in range(3):
        us = [torch.zeros_like(f) for f in fs]
        its.append(B.solve(fs, bcs, us, tol_rel=1e-12, step_max=5000))
        torch.cuda.synchronize()
        sols.append([u.cpu().numpy() for u in us])
    worst = max(np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(sols[2], sols[0]))
    print(f"[batch layout] iterations {its}; worst rel L2 vs the first layout {worst:.3e}",
          flush=True)
    assert worst <= 1e-9
    assert all(abs(i - its[0]) <= 2 + its[0] // 50 for i in its)
