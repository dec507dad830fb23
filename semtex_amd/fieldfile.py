"""semtex field files (.fld / .rst / .chk) on the host, in numpy.

The format of writeField (src/misc.cpp:106-180) and of Domain's stream
input (src/domain.cpp:329-398): per dump a 10-line header -- each value
left-justified in 25 columns, a space, then its label -- followed, for
every field letter on the "Fields written" line, by that field's nz
physical-space planes of nel*np*np doubles in the byte order the "Format"
line names.  semx_field_write / semx_field_read (include/semx.h) do the
same for the library's device fields; this module reads and writes the
files themselves, e.g. to inspect a dump or make a restart file.
"""
from __future__ import annotations

import sys
import time as _time

import numpy as np

LABELS = ("Session", "Created", "Nr, Ns, Nz, Elements", "Step", "Time", "Time step",
          "Kinvis", "Beta", "Fields written", "Format")


def header_lines(session, np_, nz, nel, step, time, dt, kinvis, beta, fields,
                 byteorder=None, created=None):
    """The ten header lines of one dump, as writeField prints them."""
    order = byteorder or sys.byteorder
    created = created or _time.strftime("%a %b %d %H:%M:%S %Y")
    vals = (f"{session:<25s}", f"{created:<25s}", f"{f'{np_} {np_} {nz} {nel}':<25s}",
            f"{step:<25d}", f"{time:<25.6g}", f"{dt:<25.6g}", f"{kinvis:<25.6g}",
            f"{beta:<25.6g}", f"{fields:<25s}", f"{'binary IEEE ' + order + '-endian':<25s}")
    return [f"{v} {lab}\n" for v, lab in zip(vals, LABELS)]


def write(path, planes, np_, nz, nel, session="session", step=0, time=0.0, dt=0.0,
          kinvis=0.0, beta=1.0, append=False, byteorder=None):
    """Write one dump: planes = {letter: array (nz, nel*np*np)} in physical
    space.  byteorder 'little' / 'big' (default: this machine's)."""
    order = byteorder or sys.byteorder
    fields = "".join(planes)
    with open(path, "ab" if append else "wb") as fh:
        for line in header_lines(session, np_, nz, nel, step, time, dt, kinvis, beta, fields,
                                 order):
            fh.write(line.encode())
        dt_ = np.dtype("<f8" if order == "little" else ">f8")
        for a in planes.values():
            a = np.asarray(a, dtype=np.float64)
            assert a.shape == (nz, nel * np_ * np_)
            fh.write(a.astype(dt_).tobytes())


def read(path):
    """All dumps of a field file: a list of dicts with the header values
    (session, created, np, nz, nel, step, time, dt, kinvis, beta, fields,
    format) and 'planes': {letter: array (nz, nel*np*np)} in native order."""
    out = []
    with open(path, "rb") as fh:
        while True:
            lines = [fh.readline() for _ in range(10)]
            if not lines[0]:
                return out
            if not all(lines):
                raise ValueError("field file: truncated header")
            txt = [ln.decode() for ln in lines]
            v = [t[:25].strip() for t in txt]
            nr, ns, nz, nel = (int(x) for x in txt[2].split()[:4])
            fields = ""
            for c in txt[8]:
                if not c.isalpha():
                    break
                fields += c
            fmt = txt[9]
            if "binary" not in fmt:
                raise ValueError("field file: not binary")
            dt_ = np.dtype(">f8" if "big" in fmt else "<f8")
            n = nz * nel * nr * ns
            planes = {}
            for c in fields:
                buf = fh.read(8 * n)
                if len(buf) != 8 * n:
                    raise ValueError("field file: truncated data")
                planes[c] = np.frombuffer(buf, dtype=dt_).astype(np.float64).reshape(nz, -1)
            out.append(dict(session=v[0], created=v[1], np=nr, nz=nz, nel=nel, step=int(v[3]),
                            time=float(v[4]), dt=float(v[5]), kinvis=float(v[6]),
                            beta=float(v[7]), fields=fields, format=fmt.strip(),
                            planes=planes))
