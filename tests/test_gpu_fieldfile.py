"""semtex field files written and read from device fields through the C ABI
(semx_field_write / semx_field_read).

- a dump holds each field in physical space (Domain::dump, src/domain.cpp:
  188-248, Nyquist plane zeroed before the inverse transform): written from
  the Fourier planes of a known physical field, the file's planes equal
  that field to rel 1e-13, and the header carries writeField's values
  (src/misc.cpp:106-180);
- a file read back gives the Fourier planes it was written from (rel
  1e-13, Nyquist plane zero) (Domain::restart, src/domain.cpp:152-180);
- a restart file made elsewhere -- big-endian, two dumps, fields in another
  order plus one not asked for -- is read by dump index with byte swapping
  (src/domain.cpp:373-397);
- a file of another plane count is refused (readField's checks,
  src/misc.cpp:196-205).
"""
import numpy as np
import pytest

import semtex_amd as sx
from semtex_amd import fieldfile as ff
from semtex_amd.mesh import rect_mesh
from semtex_amd.workloads import fourier_pack

pytestmark = pytest.mark.gpu


def physical_field(P, nz, seed):
    rng = np.random.default_rng(seed)
    z = 2 * np.pi * np.arange(nz) / nz
    a, b, c = rng.standard_normal(3)
    f = (a * np.sin(P.x[None, :] + z[:, None]) * np.cos(P.y[None, :])
         + b * np.cos(2 * z[:, None]) + c * np.sin(P.x[None, :] - 3 * z[:, None]))
    return f


@pytest.mark.parametrize("nz", [1, 2, 8, 12])
def test_write_then_read(cuda, tmp_path, nz):
    import torch
    print(f"[fieldfile] nz={nz}", flush=True)
    P = sx.Mesh(6, rect_mesh(3, 2, 0, 2 * np.pi, 0, 2 * np.pi, periodic_x=True,
                             periodic_y=True))
    phys = {c: physical_field(P, nz, s) for s, c in enumerate("uvp")}
    if nz == 2:   # two planes carry mode 0 only (AuxField::transform)
        phys = {c: np.repeat(v[:1], 2, axis=0) for c, v in phys.items()}
    dev = {}
    for c, v in phys.items():
        F = np.zeros((nz, P.planeSize))
        F[:, :P.nPlane] = fourier_pack(v) if nz > 2 else v
        if nz == 2:
            F[1] = 0.0
        dev[c] = torch.from_numpy(F.ravel().copy()).to("cuda:0")
    path = str(tmp_path / "t.fld")
    sx.write_field(path, P, nz, dev, session="tg", step=7, time=0.125, dt=0.005, kinvis=0.01,
                   beta=1.0)
    d = ff.read(path)
    assert len(d) == 1
    h = d[0]
    assert (h["session"], h["np"], h["nz"], h["nel"], h["step"], h["fields"]) == \
        ("tg", 6, nz, P.nel, 7, "uvp")
    assert (h["time"], h["dt"], h["kinvis"], h["beta"]) == (0.125, 0.005, 0.01, 1.0)
    for c, v in phys.items():
        r = np.linalg.norm(h["planes"][c] - v) / np.linalg.norm(v)
        assert r <= 1e-13, (c, r)
    back = {c: torch.zeros_like(t) for c, t in dev.items()}
    step, t, names = sx.read_field(path, P, nz, back)
    assert (step, t, names) == (7, 0.125, "uvp")
    for c in "uvp":
        a = back[c].cpu().numpy().reshape(nz, P.planeSize)
        b = dev[c].cpu().numpy().reshape(nz, P.planeSize)
        assert np.linalg.norm(a - b) <= 1e-13 * np.linalg.norm(b)
        if nz > 1:
            assert not a[1].any()


def test_read_foreign_restart(cuda, tmp_path):
    import torch
    nz = 8
    P = sx.Mesh(5, rect_mesh(2, 2, 0, 2 * np.pi, 0, 2 * np.pi, periodic_x=True,
                             periodic_y=True))
    first = {c: physical_field(P, nz, 10 + i) for i, c in enumerate("uvwp")}
    second = {c: physical_field(P, nz, 20 + i) for i, c in enumerate("pcvu")}
    path = str(tmp_path / "restart.rst")
    ff.write(path, first, P.np, nz, P.nel, session="other", step=1, byteorder="big")
    ff.write(path, second, P.np, nz, P.nel, session="other", step=2, time=0.5, append=True,
             byteorder="big")
    want = {c: torch.zeros(nz * P.planeSize, dtype=torch.float64, device="cuda:0")
            for c in "uvw"}
    before_w = want["w"].clone()
    step, t, names = sx.read_field(path, P, nz, want, dump=1)
    assert (step, t, names) == (2, 0.5, "pcvu")
    for c in "uv":
        got = want[c].cpu().numpy().reshape(nz, P.planeSize)[:, :P.nPlane]
        ref = fourier_pack(second[c])
        ref[1] = 0.0
        assert np.linalg.norm(got - ref) <= 1e-13 * np.linalg.norm(ref)
    assert torch.equal(want["w"], before_w)     # not in the second dump: left alone
    with pytest.raises(sx.SemxError):
        sx.read_field(path, P, 4, {c: torch.zeros(4 * P.planeSize, dtype=torch.float64,
                                                  device="cuda:0") for c in "u"})
