"""The all-mode Helmholtz solve captured into a HIP graph through the C ABI
(the caller's stream under torch.cuda.graph): after a first call that
allocates the solve's scratch, semx_helm_solve with iters = NULL makes no
allocation, synchronisation or memset, so it can be captured.  Every replay,
with the program's own eager work and synchronisations between replays,
must equal the eager solve bit for bit."""
import numpy as np
import pytest

import semtex_amd as sx
from solve_helpers import make_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", [sx.DIRECT, sx.MIXED], ids=["direct", "mixed"])
def test_solve_graph_replays_equal_eager(cuda, method):
    import torch
    print(f"[graph] method {method}", flush=True)
    P, Q, f, bcl, ue = make_case(8, 7, 8, 3.0, 1.0, 0.2, ("D", "N", "D", "D"))
    S = sx.ModalMatrixSys(P, 3.0, 1.0, 0, 4, 8, method)
    dev = torch.device("cuda:0")
    tf = torch.from_numpy(f.ravel().copy()).to(dev)
    tb = torch.from_numpy(bcl.ravel().copy()).to(dev)
    ue_t = torch.zeros(f.size, dtype=torch.float64, device=dev)
    sx.solve(S, tf, tb, ue_t, iters=False)
    torch.cuda.synchronize()
    ref = ue_t.cpu()
    ug = torch.zeros_like(ue_t)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        sx.solve(S, tf, tb, ug, iters=False)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        ug.zero_()
        sx.solve(S, tf, tb, ug, iters=False)
    for rep in range(4):
        ug.fill_(1.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        d = float((ug.cpu() - ref).abs().max())
        print(f"  replay {rep}: max |graph - eager| = {d:.3e}", flush=True)
        assert torch.equal(ug.cpu(), ref)
        ue_t.zero_()                        # eager work between replays
        sx.solve(S, tf, tb, ue_t, iters=False)
        torch.cuda.synchronize()
        assert torch.equal(ue_t.cpu(), ref)


def test_balanced_pcg_map_kept_after_capture(cuda, monkeypatch):
    """A MIXED solve whose PCG planes run on the workgroup map balanced by
    a first solve's iteration counts (24x24 quads np=7, Nz=16), captured
    into a graph: a later eager solve that reads the iteration counts back
    leaves the map (its buffers and grid) as the graph holds it, and every
    replay equals the eager solve bit for bit."""
    import torch
    monkeypatch.delenv("SEMX_PCG_WG", raising=False)
    print("[graph] balanced PCG map", flush=True)
    P, Q, f, bcl, ue = make_case(24, 7, 16, 0.5, 1.0, 0.0, ("D",) * 4)
    S = sx.ModalMatrixSys(P, 0.5, 1.0, 0, 8, 16, sx.MIXED)
    dev = torch.device("cuda:0")
    tf = torch.from_numpy(f.ravel().copy()).to(dev)
    tb = torch.from_numpy(bcl.ravel().copy()).to(dev)
    u = torch.zeros(f.size, dtype=torch.float64, device=dev)
    sx.solve(S, tf, tb, u)                       # counts read back: balanced map
    torch.cuda.synchronize()
    u.zero_()
    sx.solve(S, tf, tb, u, iters=False)          # runs on the balanced map
    torch.cuda.synchronize()
    w = S.sweep_info()["pcg_w"]
    ref = u.cpu()
    ug = torch.zeros_like(u)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        sx.solve(S, tf, tb, ug, iters=False)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        ug.zero_()
        sx.solve(S, tf, tb, ug, iters=False)
    for rep in range(3):
        u.zero_()
        sx.solve(S, tf, tb, u)                   # reads the counts back again
        torch.cuda.synchronize()
        assert S.sweep_info()["pcg_w"] == w
        assert torch.equal(u.cpu(), ref)
        ug.fill_(1.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        print(f"  replay {rep}: equal {torch.equal(ug.cpu(), ref)}, workgroups up to {w}",
              flush=True)
        assert torch.equal(ug.cpu(), ref)
