"""semtex's node-numbering file, session.num.

The file is what semtex's `enumerate` utility prints (printup,
utility/enumerate.cpp:426-524) and what BCmgr::buildnum reads
(src/bcmgr.cpp:586-824): a header of whitespace-separated `# KEY : value`
lines, one value column per number set (a set serves the fields named in
its column, e.g. "uvw" and "p" when their boundary conditions differ),
then one line per element-boundary node -- element (1-based), side
(1-based), offset along the side, and per set its global number (bmap)
and essential mask.  Node order is element by element, sides CCW, each side
from its first corner: the order of a NumberSys btog, which is also this
library's.
"""
from __future__ import annotations

import numpy as np

_KEYS = ("NEL", "NP_MAX", "NEXT_MAX", "NINT_MAX", "NTOTAL", "NBOUNDARY", "NGLOBAL",
         "NSOLVE", "OPTIMIZATION", "BANDWIDTH")


def read_num(path_or_text, field=None):
    """Parse a session.num file (path, or its text when it contains a
    newline).  Returns a dict per number set in file order, or the set whose
    field names contain `field`: fields, nel, np, next, nint, ntotal,
    nboundary, nglobal, nsolve, optimization, bandwidth, btog, bmask (int32
    arrays of nboundary)."""
    text = path_or_text if "\n" in path_or_text else open(path_or_text).read()
    tok = text.split()
    pos = 0

    def take(n=1):
        nonlocal pos
        out = tok[pos:pos + n]
        if len(out) < n:
            raise ValueError("session.num: file ends early")
        pos += n
        return out

    # "# FIELDS : uvwp" then a line of dashes
    if take(2) != ["#", "FIELDS"]:
        raise ValueError("session.num: expected '# FIELDS'")
    take(1)
    all_fields = take(1)[0]
    while tok[pos] != "#":
        pos += 1
    pos += 1
    while tok[pos].startswith("-"):
        pos += 1
    # "# <nset> NUMBER SETS : names"
    if tok[pos] == "#":
        pos += 1
    nset = int(take(1)[0])
    if take(3) != ["NUMBER", "SETS", ":"]:
        raise ValueError("session.num: expected 'NUMBER SETS'")
    sets = [{"fields": f} for f in take(nset)]
    for key in _KEYS:
        h, k, c = take(3)
        if h != "#" or k != key or c != ":":
            raise ValueError(f"session.num: expected '# {key} :', found {h} {k} {c}")
        for s, v in zip(sets, take(nset)):
            s[key.lower()] = int(v)
    # a line of dashes, then "# elmt side offst bmap mask ..."
    while tok[pos] == "#" or tok[pos].startswith("-"):
        pos += 1
    while not tok[pos].lstrip("-").isdigit():
        pos += 1
    nb = sets[0]["nboundary"]
    rows = np.array(tok[pos:pos + nb * (3 + 2 * nset)], dtype=np.int64)
    if rows.size != nb * (3 + 2 * nset):
        raise ValueError("session.num: fewer node lines than NBOUNDARY")
    rows = rows.reshape(nb, 3 + 2 * nset)
    for j, s in enumerate(sets):
        s["btog"] = np.ascontiguousarray(rows[:, 3 + 2 * j], dtype=np.int32)
        s["bmask"] = np.ascontiguousarray(rows[:, 4 + 2 * j], dtype=np.int32)
        s["np"] = s["np_max"]
    if field is None:
        return sets
    for s in sets:
        if field in s["fields"]:
            return s
    raise KeyError(f"session.num: no number set for field {field!r} (fields {all_fields})")


def write_num(path, mesh, fields="uvwp", optimization=1):
    """Write a mesh's numbering as session.num (one number set for all
    fields), in the layout of enumerate's printup."""
    nel, np_, nxt = mesh.nel, mesh.np, mesh.next
    vals = {"NEL": nel, "NP_MAX": np_, "NEXT_MAX": nxt, "NINT_MAX": (np_ - 2) ** 2,
            "NTOTAL": nel * np_ * np_, "NBOUNDARY": nel * nxt, "NGLOBAL": mesh.nglobal,
            "NSOLVE": mesh.nsolve, "OPTIMIZATION": optimization, "BANDWIDTH": mesh.nband}
    lines = [f"# FIELDS         :  {fields}",
             "# ----------------  -------------",
             f"# 1 NUMBER SETS  :{fields:>15}"]
    for key in _KEYS:
        lines.append(f"# {key:<15}:{vals[key]:>15}")
    lines.append("# ----------------  -------------")
    lines.append("# elmt  side offst     bmap  mask")
    nedge = np_ - 1
    i = 0
    for e in range(nel):
        for side in range(4):
            for off in range(nedge):
                lines.append(f"{e + 1:6d}{side + 1:6d}{off:6d}{int(mesh.btog[i]):9d}"
                             f"{int(mesh.bmask[i]):6d}")
                i += 1
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
