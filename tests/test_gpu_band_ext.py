"""The one-workgroup band solve over extended panels (diagonal-block
inverses folded into the panels, live rows per block) against the CPU FP64
oracle, through the C ABI, with the cooperative solve switched off: bands
up to 512 (32-column blocks, right-hand side resident in LDS), up to 1008
(16-column blocks, two panel rows per thread, LDS ring window), and wider
(the substitution sweep).  Same bar as every direct solve: rel L2 <= 1e-12."""
import os

import numpy as np
import pytest

import semtex_amd as sx
from solve_helpers import PERIODIC, drop_mode0_constant, gpu_solve, make_case, oracle_solve, rel

pytestmark = pytest.mark.gpu

CASES = [
    ("natural-2d", 4, 8, 1, 2.5, 1.0, 0.0, ("D", "N", "D", "N")),
    ("distorted-3d", 4, 6, 8, 1.0, 1.3, 0.3, ("D", "N", "D", "D")),
    ("uniform16-np9-nz64", 16, 9, 64, 20000.0, 1.0, 0.0, ("D", "D", "D", "D")),
    ("periodic-kd607-singular", 16, 9, 4, 0.0, 1.0, 0.0, PERIODIC),
    ("periodic-kd1059", 20, 12, 2, 3.0, 1.0, 0.0, PERIODIC),
]


@pytest.fixture
def no_coop():
    old = os.environ.get("SEMX_BAND_COOP")
    os.environ["SEMX_BAND_COOP"] = "0"
    yield
    if old is None:
        os.environ.pop("SEMX_BAND_COOP", None)
    else:
        os.environ["SEMX_BAND_COOP"] = old


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_ext_band_parity(cuda, no_coop, case):
    import torch
    name, nx, np_, nz, lam2, beta, dist, bc = case
    print(f"[ext band] {name}", flush=True)
    P, Q, f, bcl, ue = make_case(nx, np_, nz, lam2, beta, dist, bc)
    S = sx.ModalMatrixSys(P, lam2, beta, 0, max(nz // 2, 1), nz, sx.DIRECT)
    print(f"  band {S.info['ldab']}", flush=True)
    ug, _ = gpu_solve(torch, P, S, f, bcl)
    ug2, _ = gpu_solve(torch, P, S, f, bcl)
    uo, _ = oracle_solve(Q, nz, lam2, beta, f, bcl)
    assert np.array_equal(ug, ug2)
    if S.info["singular0"]:
        drop_mode0_constant(P, ug, uo)
    r = rel(ug[:, :P.nPlane], uo[:, :P.nPlane])
    print(f"  rel L2 gpu vs oracle = {r:.3e}", flush=True)
    assert r <= 1e-12


@pytest.mark.parametrize("nf", [2, 3])
def test_ext_band_fields_wide(cuda, no_coop, nf):
    """Several fields through one sweep of a 16-column-block factor (up to
    six right-hand sides in the LDS window) equal separate solves."""
    import torch
    print(f"[ext band fields] kd607 nf {nf}", flush=True)
    P, Q, _, _, _ = make_case(16, 9, 4, 2.0, 1.0, 0.0, PERIODIC)
    S = sx.ModalMatrixSys(P, 2.0, 1.0, 0, 2, 4, sx.DIRECT)
    cases = [make_case(16, 9, 4, 2.0, 1.0, 0.0, PERIODIC, seed=30 + i) for i in range(nf)]
    dev = torch.device("cuda:0")
    fs, bcs, us, ref = [], [], [], []
    for (_, _, f, bc, _) in cases:
        u1, _ = gpu_solve(torch, P, S, f, bc)
        ref.append(u1)
        fs.append(torch.from_numpy(f.ravel().copy()).to(dev))
        bcs.append(torch.from_numpy(bc.ravel().copy()).to(dev))
        us.append(torch.zeros(f.size, dtype=torch.float64, device=dev))
    sx.solve_fields(S, fs, bcs, us)
    torch.cuda.synchronize()
    for u, r in zip(us, ref):
        g = u.cpu().numpy().reshape(r.shape)
        d = np.abs(g[:, :P.nPlane] - r[:, :P.nPlane]).max()
        print(f"  max |batched - separate| = {d:.3e}", flush=True)
        assert d <= 1e-14 * np.abs(r).max()


_DNS_WIDTH_ORACLE = {}


def _dns_width_case(seed):
    """24x24 periodic quads, np=9, Nz=4, and its oracle solution (cached)."""
    P, Q, f, bcl, _ = make_case(24, 9, 4, 3.0, 1.0, 0.0, PERIODIC, seed=seed)
    if seed not in _DNS_WIDTH_ORACLE:
        _DNS_WIDTH_ORACLE[seed] = oracle_solve(Q, 4, 3.0, 1.0, f, bcl)[0]
    return P, f, bcl, _DNS_WIDTH_ORACLE[seed]


@pytest.mark.parametrize("pick", ["2", "3", "5", "6", "7"],
                         ids=["ring", "ring-field-split", "ring-plane-split", "ring-plane-split-1row",
                              "ring-field-split-1row"])
def test_ext_band_dns_width_candidates(cuda, no_coop, pick):
    """The band of the DNS systems (24x24 periodic quads, np=9: kd = 911,
    16-column blocks of 928 panel rows) through every ring-window candidate
    the autotuner chooses between, for one field and for three fields in one
    call, against the oracle.  The variant that ran is read back
    (semx_helm_sweep_info): a field-split candidate (3, 7) needs several
    fields, so a single field runs the sweep the library falls back to."""
    import torch
    old = os.environ.get("SEMX_AUTOTUNE")
    os.environ["SEMX_AUTOTUNE"] = pick
    try:
        print(f"[ext band dns width] SEMX_AUTOTUNE={pick}", flush=True)
        P, f, bcl, uo = _dns_width_case(40)
        S = sx.ModalMatrixSys(P, 3.0, 1.0, 0, 2, 4, sx.DIRECT)
        print(f"  band {S.info['ldab']}", flush=True)
        ug, _ = gpu_solve(torch, P, S, f, bcl)
        r = rel(ug[:, :P.nPlane], uo[:, :P.nPlane])
        ran = S.sweep_info()["pick"][1]
        print(f"  one field (variant {ran}): rel L2 gpu vs oracle = {r:.3e}", flush=True)
        assert r <= 1e-12
        if pick in ("3", "7"):
            assert ran != int(pick)
        else:
            assert ran == int(pick)
        dev = torch.device("cuda:0")
        fs, bcs, us, refs = [], [], [], []
        for seed in (41, 42, 43):
            _, f3, bc3, uo3 = _dns_width_case(seed)
            fs.append(torch.from_numpy(f3.ravel().copy()).to(dev))
            bcs.append(torch.from_numpy(bc3.ravel().copy()).to(dev))
            us.append(torch.zeros(f3.size, dtype=torch.float64, device=dev))
            refs.append(uo3)
        sx.solve_fields(S, fs, bcs, us)
        torch.cuda.synchronize()
        assert S.sweep_info()["pick"][3] == int(pick)
        for u, ref in zip(us, refs):
            g = u.cpu().numpy().reshape(ref.shape)
            r = rel(g[:, :P.nPlane], ref[:, :P.nPlane])
            print(f"  three fields: rel L2 gpu vs oracle = {r:.3e}", flush=True)
            assert r <= 1e-12
    finally:
        if old is None:
            os.environ.pop("SEMX_AUTOTUNE", None)
        else:
            os.environ["SEMX_AUTOTUNE"] = old
