"""The full-size DNS workload (Taylor-Green vortex on [0,2pi]^3 periodic,
24x24 quads np=9, Nz=96, nu=1/1600, dt=1e-3, N_TIME=3) checked through
what the flow fixes at that size -- the oracle needs minutes of setup here,
so its parity runs on smaller meshes (test_gpu_dns.py):

- kinetic energy: the vortex u = sin x cos y cos z, v = -cos x sin y cos z
  is an eigenfunction of the Laplacian (|k|^2 = 3) and the nonlinear terms
  move energy without creating any, so over 20 steps E(t)/E(0) follows
  exp(-6 nu t) (decay 7.5e-5) to within 1% of the decay (0.1% measured);
- the projected velocity is divergence-free to the splitting error:
  max |div u| <= 1e-8 of max |grad u| (1.5e-10 measured);
- the z-parity of the vortex is kept: u and v are even in z (cos z), w,
  which the nonlinear terms create, odd -- so the imaginary planes of u and
  v and the real planes of w stay at roundoff (<= 1e-12), while the
  harmonics the nonlinear terms feed (k = 0, 2, then 3, ...) grow.
Every advection form is run."""
import numpy as np
import pytest

import oracle_lib as O
import semtex_amd as sx
from semtex_amd.workloads import taylor_green_3d, workload_mesh

pytestmark = pytest.mark.gpu


def energy(F, mass, nz, n):
    """Parseval over the dDFTr packing: mode 0 and Nyquist once, others
    twice (Re and Im each)."""
    wz = np.full(nz, 2.0)
    wz[0] = wz[1] = 1.0
    return float(sum(wz[k] * np.dot(mass, F[k, :n] ** 2) for k in range(nz)))


@pytest.mark.parametrize("adv", [0, 1, 2], ids=["skew", "alternating", "convective"])
def test_taylor_green_energy_and_divergence(cuda, adv):
    import torch
    np_, m, nz = workload_mesh("periodic24_np9_nz96")
    nu, dt, beta, nstep = 1.0 / 1600, 1e-3, 1.0, 20
    print(f"[dns config3] advection {adv}, {nstep} steps", flush=True)
    P = sx.Mesh(np_, m)
    mass = O.Mesh(np_, m).factors()["Q4"]
    init = taylor_green_3d(P, nz, beta, seed=3, amp=0.0)
    ps, n = P.planeSize, P.nPlane
    e0 = sum(energy(a, mass, nz, n) for a in init)
    D = sx.DNS(P, nz, beta, dt, nu, norder=3, advection=adv, tol_rel=1e-12, step_max=500)
    f = [torch.from_numpy(a.ravel().copy()).to("cuda:0") for a in init]
    f.append(torch.zeros_like(f[0]))
    D.run(nstep, *f)
    torch.cuda.synchronize()
    U = [t.cpu().numpy().reshape(nz, ps) for t in f[:3]]
    e1 = sum(energy(a, mass, nz, n) for a in U)
    decay, want = 1.0 - e1 / e0, 1.0 - np.exp(-6.0 * nu * nstep * dt)
    # divergence: d/dx u + d/dy v + d/dz w
    div = torch.zeros_like(f[0])
    grads = []
    for i, t in enumerate(f[:3]):
        g = t.clone()
        sx.gradient(P, nz, beta, g, i)
        div += g
        grads.append(g.abs().max().item())
    torch.cuda.synchronize()
    dmax = div.abs().max().item()
    # z-parity: Im planes (2k+1, k >= 1) of u, v and Re planes (0, 2k) of w
    odd = max(np.abs(U[0][3::2, :n]).max(), np.abs(U[1][3::2, :n]).max(),
              np.abs(U[2][0::2, :n]).max())
    grown = max(np.abs(a[6:, :n]).max() for a in U)
    print(f"  energy decay {decay:.6e} (exp(-6 nu t): {want:.6e}), max|div u| {dmax:.3e} "
          f"(max|grad| {max(grads):.3f}), off-parity planes {odd:.2e}, modes >= 3 {grown:.2e}",
          flush=True)
    assert abs(decay - want) <= 1e-2 * want
    assert dmax <= 1e-8 * max(grads)
    assert odd <= 1e-12 and grown > 1e-7
