"""The mixed-precision workload at full size on the GPU: 32x32 uniform quads
on [0,2pi]^2, np=11, Nz=128 (64 modes, 127 planes), lambda2_k = k^2 (beta=1,
the pressure-like MIXED system of src/matrix.cpp:93-96), Dirichlet on every
edge, seeded smooth manufactured solution per plane.  MIXED: mode 0 by FP64
static condensation + banded Cholesky, modes 1..63 by Jacobi-PCG with the
inner iteration in FP32 (element operator in FP32, or on the BF16 matrix
cores with FP32 accumulation) and FP64 iterative refinement
(src/field.cpp:627-720 is the FP64 loop it stands in for).

Bar: rel L2 <= 1e-10 against the oracle's FP64 DIRECT solution on modes
0, 1 and 63 (three oracle factorisations of nsolve 18817 keep the test
within a minute), and every plane within 1e-8 of the manufactured solution
(discretisation error at np=11 is far below that).
"""
import numpy as np
import pytest

import oracle_lib as O
import semtex_amd as sx
from semtex_amd.workloads import helmholtz_batch, workload_mesh
from solve_helpers import rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cfg2():
    np_, m, nz = workload_mesh("square32_np11_nz128")
    print("[config2] setup", flush=True)
    P = sx.Mesh(np_, m)
    S = sx.ModalMatrixSys(P, 0.0, 1.0, 0, nz // 2, nz, sx.MIXED)
    f, bc, ue = helmholtz_batch(P, nz, 0.0, 1.0, seed=2)
    Q = O.Mesh(np_, m)
    ref = {}
    # modes 0 and 1 (planes 0, 2, 3), then mode 63 (planes 126, 127)
    for base, nmodes in ((0, 2), (63, 1)):
        k0, k1 = 2 * base, 2 * (base + nmodes)
        fo = np.ascontiguousarray(f[k0:k1].copy())
        uo = np.zeros_like(fo)
        O.field_solve_base(Q, k1 - k0, 0.0, 1.0, base, sx.DIRECT, fo,
                           np.ascontiguousarray(bc[k0:k1]), uo)
        for k in range(k0, k1):
            if k != 1:
                ref[k] = uo[k - k0, :P.nPlane].copy()
    print("[config2] oracle modes 0, 1, 63 solved", flush=True)
    return P, S, f, bc, ue, ref


@pytest.mark.parametrize("prec", [sx.PCG_FP32_REFINED, sx.PCG_BF16_REFINED],
                         ids=["fp32r", "bf16r"])
def test_config2_mixed_vs_oracle(cuda, cfg2, prec):
    import torch
    P, S, f, bc, ue, ref = cfg2
    if prec == sx.PCG_BF16_REFINED:
        S.set_pcg(prec, inner_tol=1e-2, max_outer=100)
    else:
        S.set_pcg(prec, inner_tol=1e-4, max_outer=40)
    print(f"[config2] MIXED prec {prec}", flush=True)
    dev = torch.device("cuda:0")
    tf = torch.from_numpy(f.ravel().copy()).to(dev)
    tb = torch.from_numpy(bc.ravel().copy()).to(dev)
    tu = torch.zeros_like(tf)
    it = sx.solve(S, tf, tb, tu, tol_rel=1e-13, step_max=20000)
    torch.cuda.synchronize()
    u = tu.cpu().numpy().reshape(f.shape)
    worst = max(rel(u[k, :P.nPlane], r) for k, r in ref.items())
    planes = [k for k in range(f.shape[0]) if k != 1]
    exact = float(np.abs(u[planes, :P.nPlane] - ue[planes, :P.nPlane]).max())
    print(f"  PCG iterations (max over planes) {it}, worst rel L2 vs oracle DIRECT "
          f"{worst:.3e}, max |u - u*| {exact:.3e}", flush=True)
    assert worst <= 1e-10
    assert exact <= 1e-8
