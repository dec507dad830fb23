"""DNS time steps replayed from HIP graphs (semx_dns_set_graph) against the
same steps launched kernel by kernel, through the C ABI.

Workload: the Taylor-Green vortex with a seeded 1e-3 curl perturbation on
[0,2pi]^3 periodic, 24x24 quads np=9, Nz=96, nu=1/1600, dt=1e-3, N_TIME=3,
both advection forms (the loop of dns/integrate.cpp:148-219).  The step
permutes its buffers' roles with a period of 88 steps (skew-symmetric) or
182 (alternating), so one graph is captured per role assignment.  Bar:
after the start-up steps and two full periods of replays (plus a few
steps), u, v, w and p equal the eager run's bit for bit.  Both runs force
the same band-sweep candidate, so their kernels are the same.
"""
import numpy as np
import pytest

import semtex_amd as sx
from semtex_amd.workloads import taylor_green_3d, workload_mesh

pytestmark = pytest.mark.gpu

PERIOD = {0: 88, 1: 182}


@pytest.mark.parametrize("adv", [0, 1], ids=["skew", "alternating"])
def test_dns_graph_replay_equals_eager(cuda, adv, monkeypatch):
    import torch
    monkeypatch.setenv("SEMX_AUTOTUNE", "6")
    np_, m, nz = workload_mesh("periodic24_np9_nz96")
    nu, dt, beta, norder = 1.0 / 1600, 1e-3, 1.0, 3
    P = sx.Mesh(np_, m)
    init = taylor_green_3d(P, nz, beta, seed=3)
    nsteps = norder + 2 * PERIOD[adv] + 5
    out = []
    for graphs in (False, True):
        print(f"[dns graph] advection {adv} graphs {graphs} steps {nsteps}", flush=True)
        D = sx.DNS(P, nz, beta, dt, nu, norder=norder, advection=adv, tol_rel=1e-12,
                   step_max=500)
        D.set_graph(graphs)
        fields = [torch.from_numpy(a.ravel().copy()).to("cuda:0") for a in init]
        fields.append(torch.zeros_like(fields[0]))
        # in two calls, so that the role assignment carries over between them
        D.run(norder + 1, *fields)
        D.run(nsteps - norder - 1, *fields)
        torch.cuda.synchronize()
        ng, rp = D.graph_info()
        print(f"  graphs {ng}, replayed steps {rp}", flush=True)
        if graphs:
            assert ng == PERIOD[adv]
            assert rp == nsteps - norder
        else:
            assert (ng, rp) == (0, 0)
        out.append([f.clone() for f in fields])
        del D
    for name, a, b in zip("uvwp", out[0], out[1]):
        assert torch.equal(a, b), name
    assert float(out[0][0].abs().max()) > 0.1
