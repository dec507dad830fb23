"""semtex field files on the host (CPU): the header layout of writeField
(src/misc.cpp:106-180: each value left-justified in 25 columns, a space,
the label), dumps appended one after another, and both byte orders of the
data read back (src/domain.cpp:373-397 swaps when the Format line's order
differs from the machine's)."""
import numpy as np

from semtex_amd import fieldfile as ff


def test_header_lines_layout():
    lines = ff.header_lines("taylor2", 11, 8, 4, 20, 0.4, 0.02, 0.01, np.pi, "uvwp",
                            byteorder="little", created="Thu Dec 03 09:00:00 2026")
    assert lines[0] == "taylor2                   Session\n"
    assert lines[1] == "Thu Dec 03 09:00:00 2026  Created\n"
    assert lines[2] == "11 11 8 4                 Nr, Ns, Nz, Elements\n"
    assert lines[3] == "20                        Step\n"
    assert lines[4] == "0.4                       Time\n"
    assert lines[5] == "0.02                      Time step\n"
    assert lines[6] == "0.01                      Kinvis\n"
    assert lines[7] == "3.14159                   Beta\n"
    assert lines[8] == "uvwp                      Fields written\n"
    assert lines[9] == "binary IEEE little-endian Format\n"
    assert all(ln[25] == " " for ln in lines)


def test_dumps_append_and_byte_orders(tmp_path):
    rng = np.random.default_rng(1)
    np_, nz, nel = 5, 4, 3
    a = {c: rng.standard_normal((nz, nel * np_ * np_)) for c in "uvp"}
    b = {c: rng.standard_normal((nz, nel * np_ * np_)) for c in "pu"}
    path = str(tmp_path / "run.fld")
    ff.write(path, a, np_, nz, nel, session="run", step=10, time=0.1, byteorder="big")
    ff.write(path, b, np_, nz, nel, session="run", step=20, time=0.2, append=True,
             byteorder="little")
    d = ff.read(path)
    assert [x["step"] for x in d] == [10, 20]
    assert d[0]["fields"] == "uvp" and d[1]["fields"] == "pu"
    assert "big-endian" in d[0]["format"] and "little-endian" in d[1]["format"]
    for c in "uvp":
        assert np.array_equal(d[0]["planes"][c], a[c])
    for c in "pu":
        assert np.array_equal(d[1]["planes"][c], b[c])
    raw = open(path, "rb").read()
    assert len(raw) == sum(len("".join(ff.header_lines("run", np_, nz, nel, s, t, 0.0, 0.0,
                                                         1.0, f, o, "x" * 24)).encode())
                           for s, t, f, o in ((10, 0.1, "uvp", "big"), (20, 0.2, "pu", "little"))
                           ) + 8 * nz * nel * np_ * np_ * 5
