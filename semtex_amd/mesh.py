"""Synthetic quad meshes in the form the semx C ABI (and semtex's Mesh) take.

A mesh is given element by element:
  vx, vy : four CCW vertex coordinates per element (straight sides);
  nbr    : for each element side (e*4 + s) the index e2*4 + s2 of the side
           it is joined to -- an interior neighbour or a periodic partner,
           the role of semtex's shared NODES and <P> surfaces
           (src/mesh.cpp:889-972, mesh/tgreen:49-54) -- or -1 on a boundary;
  skind  : boundary kind of each side: 0 joined, 1 essential <D>,
           2 natural <N> (src/feml.cpp SURFACES / BCS sections).
Joined sides are traversed in opposite directions by their two elements.

Element e = j*nx + i covers [x_i, x_{i+1}] x [y_j, y_{j+1}]; its sides are
0 bottom, 1 right, 2 top, 3 left, the order of ``src/mesh.cpp:772-886``.
"""
from __future__ import annotations

import numpy as np

KIND = {"D": 1, "N": 2}


def rect_mesh(nx, ny, x0=0.0, x1=1.0, y0=0.0, y1=1.0,
              bc=("D", "D", "D", "D"), periodic_x=False, periodic_y=False,
              distort=0.0):
    """Uniform nx x ny quads on [x0,x1] x [y0,y1].

    bc: kinds of the bottom, right, top and left domain sides ('D' or 'N');
    ignored in a periodic direction.  distort: amplitude (as a fraction of
    the element size) of a smooth vertex displacement that vanishes on the
    domain boundary, giving general straight-sided quads.
    """
    hx = (x1 - x0) / nx
    hy = (y1 - y0) / ny

    def vert(ii, jj):
        x = x0 + hx * ii
        y = y0 + hy * jj
        if distort:
            sx = np.sin(np.pi * ii / nx) * np.sin(2 * np.pi * jj / ny)
            sy = np.sin(2 * np.pi * ii / nx) * np.sin(np.pi * jj / ny)
            x += distort * hx * sx
            y += distort * hy * sy
        return x, y

    nel = nx * ny
    vx = np.zeros((nel, 4))
    vy = np.zeros((nel, 4))
    nbr = -np.ones((nel, 4), dtype=np.int32)
    skind = np.zeros((nel, 4), dtype=np.int32)
    corners = ((0, 0), (1, 0), (1, 1), (0, 1))
    for j in range(ny):
        for i in range(nx):
            e = j * nx + i
            for c, (di, dj) in enumerate(corners):
                vx[e, c], vy[e, c] = vert(i + di, j + dj)
            # bottom (0) <-> top (2) of the element below; right (1) <-> left (3)
            if j > 0 or periodic_y:
                nbr[e, 0] = ((j - 1) % ny * nx + i) * 4 + 2
            else:
                skind[e, 0] = KIND[bc[0]]
            if j < ny - 1 or periodic_y:
                nbr[e, 2] = ((j + 1) % ny * nx + i) * 4 + 0
            else:
                skind[e, 2] = KIND[bc[2]]
            if i < nx - 1 or periodic_x:
                nbr[e, 1] = (j * nx + (i + 1) % nx) * 4 + 3
            else:
                skind[e, 1] = KIND[bc[1]]
            if i > 0 or periodic_x:
                nbr[e, 3] = (j * nx + (i - 1) % nx) * 4 + 1
            else:
                skind[e, 3] = KIND[bc[3]]
    return dict(vx=vx, vy=vy, nbr=nbr, skind=skind, nel=nel)


def side_nodes(np_, side):
    """Row-major element-node indices of one side, CCW order
    (``src/mesh.cpp:792-850``)."""
    nm = np_ - 1
    if side == 0:
        return [i for i in range(np_)]
    if side == 1:
        return [i * np_ + nm for i in range(np_)]
    if side == 2:
        return [nm * np_ + (nm - i) for i in range(np_)]
    return [(nm - i) * np_ for i in range(np_)]


def bc_lines(field, belmt, bside, np_):
    """BC line values (nbound*np per plane) taken from a nodal field of shape
    (nz, nel*np*np) along each boundary side, in the boundary-list order."""
    field = np.asarray(field)
    nz = field.shape[0]
    npnp = np_ * np_
    out = np.zeros((nz, max(len(belmt), 0) * np_))
    for b, (e, s) in enumerate(zip(belmt, bside)):
        idx = [e * npnp + k for k in side_nodes(np_, s)]
        out[:, b * np_:(b + 1) * np_] = field[:, idx]
    return out


def node_mesh(xy, elements, side_kind=None, default_kind="D"):
    """A mesh from semtex-style NODES and ELEMENTS (mesh/kovas1: <NODES>
    coordinates, <ELEMENTS> four CCW node numbers each, 0-based here).
    Sides shared by two elements are joined; every other side is a boundary
    of kind side_kind.get((e, s), default_kind) ('D' or 'N'), the role of
    the session's <SURFACES> and <BCS>."""
    xy = np.asarray(xy, dtype=np.float64)
    elements = np.asarray(elements, dtype=np.int64)
    nel = len(elements)
    vx = xy[elements, 0].copy()
    vy = xy[elements, 1].copy()
    nbr = -np.ones((nel, 4), dtype=np.int32)
    skind = np.zeros((nel, 4), dtype=np.int32)
    owner = {}
    for e in range(nel):
        for s in range(4):
            a, b = int(elements[e, s]), int(elements[e, (s + 1) % 4])
            key = (min(a, b), max(a, b))
            if key in owner:
                e2, s2 = owner.pop(key)
                nbr[e, s] = e2 * 4 + s2
                nbr[e2, s2] = e * 4 + s
            else:
                owner[key] = (e, s)
    for (e, s) in owner.values():
        kind = (side_kind or {}).get((e, s), default_kind)
        skind[e, s] = KIND[kind]
    return dict(vx=vx, vy=vy, nbr=nbr, skind=skind, nel=nel)
