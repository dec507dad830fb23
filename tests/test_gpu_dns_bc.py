"""The DNS time step on meshes with boundaries (semx_dns_create_bc): the
computed high-order pressure flux (BCmgr::maintainFourier and
evaluateCNBCp, src/bcmgr.cpp:1039-1337) with given velocity BC lines,
through the C ABI.

Bars (FP64):
- kovas1 (mesh/kovas1, regress/kovas1.ok): the error norms of u, v and p
  after 400 steps to every printed digit, with both meshes numbered as
  `enumerate -O3` numbers them (semx_mesh_create_enumerated; the numbering
  a session.num of the regression run holds), and rel L2 <= 1e-12 against
  the oracle's run;
- a sheared channel with slanted walls and an outflow side (velocity
  natural, pressure essential there), 2D and 3D (nz = 8), N_TIME 2 and 3,
  both advection forms: rel L2 <= 1e-12 against the oracle, every field
  and plane but the Nyquist plane, with both solvers on one numbering --
  after 8 steps from the BC field itself, and after 3 steps from a shifted
  field (u.n changes on the boundary, so the flux's time-derivative term
  is live; compared at the transient, where the pressure is large, since
  roundoff of order 1e-16 of the transient's scale remains once it has
  decayed).  dt = 1e-2: the pressure is formed from terms of order |u|/dt
  and is much smaller than they are, so at dt = 2e-3 the oracle's own
  pressure moves by 2e-12 when its inputs are perturbed by 1e-15; at
  1e-2 by at most 2.4e-13;
- the same steps replayed from HIP graphs equal the eager steps bit for
  bit (the flux history lives in fixed arrays, so the graphs stay valid).
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
import semtex_amd as sx
from bounded_cases import (KOVAS, KOVAS_OK, error_norms, fourier_planes, kovas_fields,
                           kovas_meshes, sheared_bc_lines, sheared_meshes)
from semtex_amd.workloads import fourier_pack
from solve_helpers import oracle_numbering, rel

pytestmark = pytest.mark.gpu


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64).ravel().copy()).to("cuda:0")


def host(t, shape):
    return t.cpu().numpy().reshape(shape)


def test_kovas1_regress(cuda):
    import torch
    print("[dns bc] kovas1, 400 steps", flush=True)
    mu, mp = kovas_meshes()
    Qu, Qp = O.Mesh(KOVAS["np"], mu), O.Mesh(KOVAS["np"], mp)
    Qu.enumerate(3)
    Qp.enumerate(3)
    # the library's own enumerate -O3 numbering (equal to the oracle's,
    # checked on the CPU)
    Pu = sx.Mesh(KOVAS["np"], mu, numbering=("enumerate", 3))
    Pp = sx.Mesh(KOVAS["np"], mp, numbering=("enumerate", 3))
    (u, v, p), (bcu, bcv) = kovas_fields(Pu)
    nu, dt = 1.0 / KOVAS["Re"], KOVAS["dt"]
    tu, tv, tp = (dev(torch, a) for a in (u, v, p))
    D = sx.DNS(Pu, 1, 1.0, dt, nu, norder=KOVAS["norder"], advection=KOVAS["advection"],
               tol_rel=1e-12, step_max=500, pressure_mesh=Pp,
               bc=(dev(torch, bcu), dev(torch, bcv)))
    D.run(KOVAS["nstep"], tu, tv, None, tp)
    torch.cuda.synchronize()
    assert D.steps == KOVAS["nstep"]
    shape = (1, Pu.planeSize)
    gu, gv, gp = host(tu, shape), host(tv, shape), host(tp, shape)
    got = error_norms(Pu, gu, gv, gp)
    print(f"  norms {got} (regress/kovas1.ok {KOVAS_OK})", flush=True)
    assert got == KOVAS_OK
    # the oracle's own run on the same numbering
    Do = O.lib().orc_dns_create_bc(C.c_void_p(Qu.h), C.c_void_p(Qp.h), 1, 1.0, dt, nu,
                                   KOVAS["norder"], KOVAS["advection"], 1e-12, 500, 0,
                                   O.dp(bcu), O.dp(bcv), None)
    t = C.c_double()
    O.lib().orc_dns_run(C.c_void_p(Do), KOVAS["nstep"], O.dp(u), O.dp(v), None, O.dp(p),
                        C.byref(t))
    O.lib().orc_dns_destroy(C.c_void_p(Do))
    n = Pu.nPlane
    for name, g, o in (("u", gu, u), ("v", gv, v), ("p", gp, p)):
        r = rel(g[:, :n], o[:, :n])
        print(f"  {name}: rel L2 vs oracle {r:.3e}", flush=True)
        assert r <= 1e-12


def sheared_case(nz, shift, np_=7):
    mu, mp = sheared_meshes()
    Qu, Qp = O.Mesh(np_, mu), O.Mesh(np_, mp)
    Pu = sx.Mesh(np_, mu, numbering=oracle_numbering(Qu))
    Pp = sx.Mesh(np_, mp, numbering=oracle_numbering(Qp))
    lines = sheared_bc_lines(Pu, fourier_planes(Pu, nz, 0.0, fourier_pack))
    init = fourier_planes(Pu, nz, shift, fourier_pack)
    return Qu, Qp, Pu, Pp, lines, init


def run_gpu(torch, Pu, Pp, lines, init, nz, norder, adv, nstep, dt, nu, graphs=False,
            split=None):
    ncom = 3 if nz > 1 else 2
    D = sx.DNS(Pu, nz, 1.0, dt, nu, norder=norder, advection=adv, tol_rel=1e-14,
               step_max=2000, pressure_mesh=Pp, bc=[dev(torch, a) for a in lines[:ncom]])
    D.set_graph(graphs)
    f = [dev(torch, a) for a in init[:ncom]]
    if ncom == 2:
        f.append(None)
    f.append(torch.zeros_like(f[0]))
    for k in (split or [nstep]):
        D.run(k, *f)
    torch.cuda.synchronize()
    return D, f


SHEARED = [(1, 3, 1, 0.0, 8), (8, 2, 0, 0.0, 8), (8, 3, 1, 0.0, 8),
           (1, 3, 1, 0.1, 3), (8, 2, 0, 0.1, 3)]


@pytest.mark.parametrize("nz,norder,adv,shift,nstep", SHEARED,
                         ids=["2d-o3-alt", "3d-o2-skew", "3d-o3-alt", "2d-o3-alt-transient",
                              "3d-o2-skew-transient"])
def test_sheared_channel_matches_oracle(cuda, nz, norder, adv, shift, nstep):
    import torch
    print(f"[dns bc] sheared channel nz={nz} N_TIME={norder} advection={adv} shift={shift} "
          f"steps={nstep}", flush=True)
    Qu, Qp, Pu, Pp, lines, init = sheared_case(nz, shift)
    dt, nu = 1e-2, 0.05
    D, f = run_gpu(torch, Pu, Pp, lines, init, nz, norder, adv, nstep, dt, nu)
    assert D.steps == nstep
    ncom = 3 if nz > 1 else 2
    o = [np.ascontiguousarray(a.copy()) for a in init[:ncom]]
    po = np.zeros_like(o[0])
    Do = O.lib().orc_dns_create_bc(C.c_void_p(Qu.h), C.c_void_p(Qp.h), nz, 1.0, dt, nu, norder,
                                   adv, 1e-14, 2000, 0, *[O.dp(a) for a in lines[:2]],
                                   O.dp(lines[2]) if ncom == 3 else None)
    t = C.c_double()
    O.lib().orc_dns_run(C.c_void_p(Do), nstep, O.dp(o[0]), O.dp(o[1]),
                        O.dp(o[2]) if ncom == 3 else None, O.dp(po), C.byref(t))
    O.lib().orc_dns_destroy(C.c_void_p(Do))
    keep = [k for k in range(nz) if not (nz > 1 and k == 1)]
    shape = (nz, Pu.planeSize)
    gots = [f[i] for i in range(ncom)] + [f[3]]
    for name, g, oo in zip("uvw"[:ncom] + "p", gots, o + [po]):
        a = host(g, shape)[keep, :Pu.nPlane]
        b = oo[keep, :Pu.nPlane]
        r = rel(a, b)
        print(f"  {name}: rel L2 vs oracle {r:.3e}", flush=True)
        assert r <= 1e-12
        assert np.abs(b).max() > 1e-3


def test_bounded_graph_replay_equals_eager(cuda, monkeypatch):
    import torch
    monkeypatch.setenv("SEMX_AUTOTUNE", "6")
    nz, norder, adv, nstep = 8, 3, 0, 40
    _, _, Pu, Pp, lines, init = sheared_case(nz, 0.1)
    out = []
    for graphs in (False, True):
        print(f"[dns bc graph] graphs {graphs} steps {nstep}", flush=True)
        D, f = run_gpu(torch, Pu, Pp, lines, init, nz, norder, adv, nstep, 2e-3, 0.05,
                       graphs=graphs, split=[norder + 1, nstep - norder - 1])
        ng, rp = D.graph_info()
        print(f"  graphs {ng}, replayed steps {rp}", flush=True)
        assert rp == (nstep - norder if graphs else 0)
        out.append([t.clone() for t in f])
        del D
    for name, a, b in zip("uvwp", out[0], out[1]):
        assert torch.equal(a, b), name


def test_create_bc_rejects_mismatched_meshes(cuda):
    import torch
    mu, _ = sheared_meshes()
    other = sheared_meshes(5, 3)[1]
    Pu, Po = sx.Mesh(5, mu), sx.Mesh(5, other)
    nl = Pu.nbound * Pu.np
    z = torch.zeros(nl, dtype=torch.float64, device="cuda:0")
    with pytest.raises(sx.SemxError):
        sx.DNS(Pu, 1, 1.0, 1e-3, 0.1, pressure_mesh=Po, bc=(z, z))


def test_time_dependent_bcs_match_oracle(cuda):
    """semx_dns_set_bc before each step (the decaying Taylor vortex in a box,
    walls moving with the decay): rel L2 <= 1e-12 against the oracle's same
    sequence, which ends near the exact solution (tests/test_oracle.py)."""
    import torch
    from bounded_cases import BOX, box_lines, box_meshes
    from test_oracle import box_run_oracle
    print(f"[dns bc] time-dependent BCs, {BOX['nstep']} steps", flush=True)
    mu, mp = box_meshes()
    Q = O.Mesh(BOX["np"], mu)
    Qp = O.Mesh(BOX["np"], mp)
    Pu = sx.Mesh(BOX["np"], mu, numbering=oracle_numbering(Q))
    Pp = sx.Mesh(BOX["np"], mp, numbering=oracle_numbering(Qp))
    (u, v), (bu, bv) = box_lines(Pu, 0.0, BOX["nu"])
    D = sx.DNS(Pu, 1, 1.0, BOX["dt"], BOX["nu"], norder=BOX["norder"], advection=1,
               tol_rel=1e-12, step_max=500, pressure_mesh=Pp, bc=(dev(torch, bu), dev(torch, bv)))
    tu, tv = dev(torch, u), dev(torch, v)
    tp = torch.zeros_like(tu)
    for n in range(BOX["nstep"]):
        _, (bu, bv) = box_lines(Pu, (n + 1) * BOX["dt"], BOX["nu"])
        D.set_bc([dev(torch, bu), dev(torch, bv)])
        D.run(1, tu, tv, None, tp)
    torch.cuda.synchronize()
    _, uo, vo = box_run_oracle(True)
    n = Pu.nPlane
    for name, g, o in (("u", tu, uo), ("v", tv, vo)):
        r = rel(host(g, (1, Pu.planeSize))[:, :n], o[:, :n])
        print(f"  {name}: rel L2 vs oracle {r:.3e}", flush=True)
        assert r <= 1e-12


def test_sheared_channel_iterative_matches_oracle(cuda):
    """ITERATIVE 2 (velocity JACPCG, pressure MIXED) on the sheared 3D
    channel with its transient: rel L2 <= 1e-10 against the oracle's run
    at the same level, every field and plane but the Nyquist plane."""
    import torch
    nz, norder, adv, shift, nstep = 8, 3, 1, 0.1, 3
    print(f"[dns bc] sheared channel ITERATIVE 2 nz={nz} steps={nstep}", flush=True)
    Qu, Qp, Pu, Pp, lines, init = sheared_case(nz, shift)
    dt, nu = 1e-2, 0.05
    D = sx.DNS(Pu, nz, 1.0, dt, nu, norder=norder, advection=adv, tol_rel=1e-14,
               step_max=5000, pressure_mesh=Pp, bc=[dev(torch, a) for a in lines])
    D.set_iterative(2)
    f = [dev(torch, a) for a in init] + [torch.zeros(nz * Pu.planeSize, dtype=torch.float64,
                                                     device="cuda:0")]
    D.run(nstep, *f)
    torch.cuda.synchronize()
    o = [np.ascontiguousarray(a.copy()) for a in init]
    po = np.zeros_like(o[0])
    Do = O.lib().orc_dns_create_bc(C.c_void_p(Qu.h), C.c_void_p(Qp.h), nz, 1.0, dt, nu, norder,
                                   adv, 1e-14, 5000, 0, *[O.dp(a) for a in lines])
    assert O.lib().orc_dns_set_iterative(C.c_void_p(Do), 2) == 0
    t = C.c_double()
    O.lib().orc_dns_run(C.c_void_p(Do), nstep, *[O.dp(a) for a in o], O.dp(po), C.byref(t))
    O.lib().orc_dns_destroy(C.c_void_p(Do))
    keep = [k for k in range(nz) if k != 1]
    for name, g, oo in zip("uvwp", f, o + [po]):
        r = rel(host(g, (nz, Pu.planeSize))[keep, :Pu.nPlane], oo[keep, :Pu.nPlane])
        print(f"  {name}: rel L2 vs oracle {r:.3e}", flush=True)
        assert r <= 1e-10
