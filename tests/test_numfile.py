"""semtex's numbering at the boundary, checked on the CPU (no GPU calls).

- session.num (utility/enumerate.cpp:426-524 writes it, src/bcmgr.cpp:
  586-824 reads it): a file in enumerate's layout, with two number sets,
  parses to the values it holds; the library's own numbering written in
  that layout reads back unchanged;
- semx_mesh_create_numbered takes a numbering other than the library's own
  (reverse order of the unknowns) and keeps it, with the band recomputed by
  the formula of utility/enumerate.cpp:903-943; the oracle adopts the same
  numbering;
- numberings that do not fit the mesh are refused (SEMX_E_MESH), as
  buildnum refuses a mismatched file.
"""
import numpy as np
import pytest

import oracle_lib as O
import semtex_amd as sx
from semtex_amd.mesh import rect_mesh
from semtex_amd.numfile import read_num, write_num

SEMX_E_MESH = 7

# enumerate's printup layout for a 1-element np=3 mesh with two number sets
# ("uvw" essential on side 1, "p" natural everywhere)
TWO_SETS = """# FIELDS         :  uvwp
# ----------------  -------------  -------------
# 2 NUMBER SETS  :            uvw              p
# NEL            :              1              1
# NP_MAX         :              3              3
# NEXT_MAX       :              8              8
# NINT_MAX       :              1              1
# NTOTAL         :              9              9
# NBOUNDARY      :              8              8
# NGLOBAL        :              8              8
# NSOLVE         :              5              8
# OPTIMIZATION   :              1              1
# BANDWIDTH      :              5              8
# ----------------  -------------  -------------
# elmt  side offst     bmap  mask     bmap  mask
     1     1     0        5     1        0     0
     1     1     1        6     1        1     0
     1     2     0        7     1        2     0
     1     2     1        0     0        3     0
     1     3     0        1     0        4     0
     1     3     1        2     0        5     0
     1     4     0        3     0        6     0
     1     4     1        4     0        7     0
"""


def band(btog, bmask, nel, nxt):
    nb = 0
    for e in range(nel):
        g = [btog[e * nxt + i] for i in range(nxt) if not bmask[e * nxt + i]]
        if g:
            nb = max(nb, max(g) - min(g))
    return nb + 1


def test_read_two_number_sets():
    sets = read_num(TWO_SETS)
    assert [s["fields"] for s in sets] == ["uvw", "p"]
    u, p = sets
    assert (u["nel"], u["np"], u["nglobal"], u["nsolve"], u["bandwidth"]) == (1, 3, 8, 5, 5)
    assert (p["nsolve"], p["bandwidth"]) == (8, 8)
    assert list(u["btog"]) == [5, 6, 7, 0, 1, 2, 3, 4]
    assert list(u["bmask"]) == [1, 1, 1, 0, 0, 0, 0, 0]
    assert list(p["btog"]) == list(range(8))
    assert read_num(TWO_SETS, field="p")["fields"] == "p"
    assert read_num(TWO_SETS, field="w")["fields"] == "uvw"
    with pytest.raises(KeyError):
        read_num(TWO_SETS, field="c")


@pytest.mark.parametrize("kw", [dict(nx=5, ny=4, bc=("D", "N", "D", "D")),
                                dict(nx=4, ny=4, periodic_x=True, periodic_y=True)],
                         ids=["dirichlet-natural", "periodic"])
def test_write_read_round_trip(tmp_path, kw):
    P = sx.Mesh(6, rect_mesh(**kw))
    path = str(tmp_path / "session.num")
    write_num(path, P, fields="uvwp")
    s = read_num(path, field="u")
    assert s["fields"] == "uvwp"
    assert (s["nel"], s["np"], s["next_max"], s["nint_max"]) == (P.nel, P.np, P.next, P.nint)
    assert (s["nglobal"], s["nsolve"], s["bandwidth"]) == (P.nglobal, P.nsolve, P.nband)
    assert s["ntotal"] == P.nel * P.np * P.np and s["nboundary"] == P.nel * P.next
    assert np.array_equal(s["btog"], P.btog) and np.array_equal(s["bmask"], P.bmask)


def reversed_numbering(P):
    """Another valid numbering: unknowns in reverse order, essential last."""
    bt = P.btog.copy()
    unk = bt < P.nsolve
    bt[unk] = P.nsolve - 1 - bt[unk]
    return dict(btog=bt, bmask=P.bmask.copy(), nglobal=P.nglobal, nsolve=P.nsolve)


@pytest.mark.parametrize("kw", [dict(nx=5, ny=4, bc=("D", "N", "D", "D"), distort=0.2),
                                dict(nx=4, ny=3, periodic_x=True, periodic_y=True)],
                         ids=["dirichlet-natural", "periodic"])
def test_numbered_mesh_keeps_given_numbering(tmp_path, kw):
    m = rect_mesh(**kw)
    P = sx.Mesh(7, m)
    num = reversed_numbering(P)
    path = str(tmp_path / "session.num")
    Pn = sx.Mesh(7, m, numbering=num)
    write_num(path, Pn)
    back = read_num(path, field="p")
    assert np.array_equal(Pn.btog, num["btog"]) and np.array_equal(Pn.bmask, num["bmask"])
    assert np.array_equal(back["btog"], num["btog"])
    assert Pn.nband == band(num["btog"], num["bmask"], P.nel, P.next)
    assert (Pn.nglobal, Pn.nsolve) == (P.nglobal, P.nsolve)
    Q = O.Mesh(7, m)
    Q.renumber(num)
    assert np.array_equal(Q.btog, num["btog"]) and Q.nband == Pn.nband


def test_unfit_numberings_refused():
    import ctypes as C
    m = rect_mesh(4, 3, bc=("D", "N", "D", "D"))
    P = sx.Mesh(5, m)
    good = reversed_numbering(P)
    L = sx.load()

    def create(num):
        vx = np.ascontiguousarray(m["vx"], dtype=np.float64)
        vy = np.ascontiguousarray(m["vy"], dtype=np.float64)
        nbr = np.ascontiguousarray(m["nbr"], dtype=np.int32)
        sk = np.ascontiguousarray(m["skind"], dtype=np.int32)
        bt = np.ascontiguousarray(num["btog"], dtype=np.int32)
        bm = np.ascontiguousarray(num["bmask"], dtype=np.int32)
        h = C.c_void_p()
        st = L.semx_mesh_create_numbered(5, int(m["nel"]), sx._dp(vx), sx._dp(vy), sx._ip(nbr),
                                         sx._ip(sk), -1, 0.0, 1.0, sx._ip(bt), sx._ip(bm),
                                         int(num["nglobal"]), int(num["nsolve"]), C.byref(h))
        if st == 0:
            L.semx_mesh_destroy(h)
        return st

    assert create(good) == 0
    # two slots of one joined node given different numbers
    bad = dict(good, btog=good["btog"].copy())
    i = int(np.flatnonzero(np.bincount(bad["btog"]) > 1)[0])
    slots = np.flatnonzero(bad["btog"] == i)
    j = int(np.flatnonzero((bad["btog"] != i) & (bad["bmask"] == bad["bmask"][slots[0]]))[0])
    bad["btog"][slots[0]] = bad["btog"][j]
    assert create(bad) == SEMX_E_MESH
    # an essential node numbered among the unknowns
    swap = dict(good, btog=good["btog"].copy())
    e_id = int(swap["btog"][swap["bmask"] == 1][0])
    swap["btog"][swap["btog"] == e_id] = -2
    swap["btog"][swap["btog"] == 0] = e_id
    swap["btog"][swap["btog"] == -2] = 0
    assert create(swap) == SEMX_E_MESH
    # mask that disagrees with the boundary conditions
    mask = dict(good, bmask=good["bmask"].copy())
    mask["bmask"][np.flatnonzero(mask["bmask"] == 1)[0]] = 0
    assert create(mask) == SEMX_E_MESH
    # wrong counts
    assert create(dict(good, nsolve=good["nsolve"] - 1)) == SEMX_E_MESH
    with pytest.raises(ValueError):
        O.Mesh(5, m).renumber(bad)
