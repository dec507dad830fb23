"""The cooperative band solve (several workgroups per mode in a lagged
pipeline) against the CPU FP64 oracle, through the C ABI: same bar as the
one-workgroup sweep, rel L2 <= 1e-12, over lane layouts from the default
to one chunk per lane."""
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
# (name, lead chunks, follower chunks, extended panels): the cooperative
# sweep over separate diagonal-block inverses, and the one over extended
# panels (SEMX_BAND_XCOOP; bands up to 1008, wider ones fall back to the
# former)
LAYOUTS = [("default", None, None, "0"), ("one-chunk-lanes", "1", "1", "0"),
           ("lead2-follow3", "2", "3", "0"), ("xcoop-default", None, None, "1"),
           ("xcoop-one-chunk-lanes", "1", "1", "1"), ("xcoop-lead3-follow2", "3", "2", "1")]


@pytest.fixture
def coop_env(request):
    name, lead, follow, xcoop = request.param
    keys = ("SEMX_BAND_COOP", "SEMX_COOP_LEAD", "SEMX_COOP_FOLLOW", "SEMX_BAND_XCOOP")
    old = {k: os.environ.get(k) for k in keys}
    os.environ["SEMX_BAND_COOP"] = "1"
    os.environ["SEMX_BAND_XCOOP"] = xcoop
    for k, v in (("SEMX_COOP_LEAD", lead), ("SEMX_COOP_FOLLOW", follow)):
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v
    yield name
    for k, v in old.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v


@pytest.mark.parametrize("coop_env", LAYOUTS, ids=[l[0] for l in LAYOUTS], indirect=True)
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_coop_band_parity(cuda, coop_env, case):
    import torch
    name, nx, np_, nz, lam2, beta, dist, bc = case
    print(f"[coop band] {name} layout {coop_env}", flush=True)
    P, Q, f, bcl, ue = make_case(nx, np_, nz, lam2, beta, dist, bc)
    S = sx.ModalMatrixSys(P, lam2, beta, 0, max(nz // 2, 1), nz, sx.DIRECT)
    ug, _ = gpu_solve(torch, P, S, f, bcl)
    ug2, _ = gpu_solve(torch, P, S, f, bcl)     # a second launch (next flag epoch)
    uo, _ = oracle_solve(Q, nz, lam2, beta, f, bcl)
    assert np.array_equal(ug, ug2)
    # an asynchronous solve (iters = NULL) reports through the next check:
    # no hand-off wait of these launches expired
    tf = torch.from_numpy(np.ascontiguousarray(f).ravel().copy()).to("cuda:0")
    tb = torch.from_numpy(np.ascontiguousarray(bcl).ravel().copy()).to("cuda:0")
    tu = torch.zeros_like(tf)
    sx.solve(S, tf, tb, tu, iters=False)
    sx.helm_check(S)
    assert np.array_equal(tu.cpu().numpy().reshape(f.shape), ug)
    if S.info["singular0"]:
        drop_mode0_constant(P, ug, uo)
    r = rel(ug[:, :P.nPlane], uo[:, :P.nPlane])
    print(f"  rel L2 gpu vs oracle = {r:.3e}", flush=True)
    assert r <= 1e-12


@pytest.mark.parametrize("coop_env", [LAYOUTS[0], LAYOUTS[3]], ids=["default", "xcoop"],
                         indirect=True)
def test_coop_band_fields(cuda, coop_env):
    """Three fields sharing the system in one cooperative sweep."""
    import torch
    P, Q, _, _, _ = make_case(8, 7, 8, 3.0, 1.0, 0.2, ("D", "N", "D", "D"))
    S = sx.ModalMatrixSys(P, 3.0, 1.0, 0, 4, 8, sx.DIRECT)
    cases = [make_case(8, 7, 8, 3.0, 1.0, 0.2, ("D", "N", "D", "D"), seed=30 + i)
             for i in range(3)]
    dev = torch.device("cuda:0")
    fs = [torch.from_numpy(c[2].ravel().copy()).to(dev) for c in cases]
    bcs = [torch.from_numpy(c[3].ravel().copy()).to(dev) for c in cases]
    us = [torch.zeros(c[2].size, dtype=torch.float64, device=dev) for c in cases]
    sx.solve_fields(S, fs, bcs, us)
    torch.cuda.synchronize()
    for c, u in zip(cases, us):
        uo, _ = oracle_solve(Q, 8, 3.0, 1.0, c[2], c[3])
        g = u.cpu().numpy().reshape(c[2].shape)
        r = rel(g[:, :P.nPlane], uo[:, :P.nPlane])
        print(f"  field rel L2 vs oracle = {r:.3e}", flush=True)
        assert r <= 1e-12
