"""GPU parity with spectral vanishing viscosity switched on, through the C
ABI, against the CPU FP64 oracle with the same SVV settings.

- x-y: SVV_MN, SVV_EPSN, KINVIS at mesh creation select the SVV derivative
  SDV = BW diag(sqrt S) FW DV (src/svv.cpp:126-185) wherever lambda2 >
  EPSDP (src/element.cpp:1484,1532,1640): condensed matrices, band factor,
  the PCG operator;
- z: SVV_MZ, SVV_EPSZ weight betak2 of the modes above SVV_MZ
  (SVV::coeffs_z, src/svv.cpp:82-123, in ModalMatrixSys, src/matrix.cpp:
  78-83).

Bars: rel L2 <= 1e-12 vs the oracle (DIRECT), <= 1e-10 (MIXED: mode 0
direct, the others JACPCG to 1e-13).  A solve with SVV off differs from
the SVV solve by far more than the bar, so the settings are exercised.
"""
import numpy as np
import pytest

import oracle_lib as O
import semtex_amd as sx
from semtex_amd.mesh import rect_mesh
from semtex_amd.workloads import helmholtz_batch
from solve_helpers import gpu_solve, oracle_solve, rel, solved

pytestmark = pytest.mark.gpu

SVV = dict(svv_mn=4, svv_eps=0.05, kinvis=0.01)       # S_N = 6 at np = 9
SVVZ = dict(svv_mz=1, svv_epsz=0.02, kinvis=0.01)     # S_z = 3 at the last mode


@pytest.mark.parametrize("method", [sx.DIRECT, sx.MIXED], ids=["direct", "mixed"])
def test_svv_xy_and_z_vs_oracle(cuda, method):
    import torch
    np_, nz, lam2, beta = 9, 8, 4.0, 1.0
    m = rect_mesh(4, 4, 0, 2 * np.pi, 0, 2 * np.pi, distort=0.15, bc=("D", "N", "D", "D"))
    print(f"[svv] method {method}", flush=True)
    P = sx.Mesh(np_, m, **SVV)
    Q = O.Mesh(np_, m, **SVV)
    f, bc, _ = helmholtz_batch(P, nz, lam2, beta, seed=6)
    S = sx.ModalMatrixSys(P, lam2, beta, 0, nz // 2, nz, method, nmode_total=nz // 2, **SVVZ)
    if method == sx.MIXED:
        S.set_pcg(sx.PCG_FP64)
    ug, _ = gpu_solve(torch, P, S, f, bc, tol=1e-13)
    O.set_svv_z(SVVZ["svv_mz"], SVVZ["svv_epsz"], SVVZ["kinvis"], nz // 2)
    try:
        uo, _ = oracle_solve(Q, nz, lam2, beta, f, bc, method=method, tol=1e-13)
    finally:
        O.set_svv_z()
    r = rel(solved(P, ug), solved(P, uo))
    # the same problem without SVV: the settings change the answer
    P0 = sx.Mesh(np_, m)
    S0 = sx.ModalMatrixSys(P0, lam2, beta, 0, nz // 2, nz, sx.DIRECT)
    u0, _ = gpu_solve(torch, P0, S0, f, bc)
    d = rel(solved(P, u0), solved(P, uo))
    print(f"  rel L2 vs oracle {r:.3e}; SVV off differs by {d:.3e}", flush=True)
    assert r <= (1e-12 if method == sx.DIRECT else 1e-10)
    assert d > 1e-6
