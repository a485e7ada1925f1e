"""Workload meshes as plain arrays (no reference counterpart: the reference
generates only uniform boxes, io/generators.cpp:10-103).

A mesh here is the reference's PolyMesh layout (mesh/mesh.hpp:36-68): FP64
points, faces as node lists (an area vector by the right-hand rule points from
owner to neighbour, or out of the domain on boundary faces), internal faces
first, boundary faces grouped into contiguous named patches. Both the HIP
solver and the oracle renumber it (reverse Cuthill-McKee) on construction, so
the input order of cells and faces does not reach the results.

structured_hex builds any union of cells of a tensor-product grid, which
covers graded boxes and the backward-facing step.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class PolyMeshArrays:
    points: np.ndarray        # (n_points, 3) float64
    face_offsets: np.ndarray  # (n_faces + 1,) int32
    face_nodes: np.ndarray    # int32
    owner: np.ndarray         # (n_faces,) int32
    neighbour: np.ndarray     # (n_internal,) int32
    n_cells: int
    patches: list = field(default_factory=list)  # [(name, start, size)]

    @property
    def n_faces(self) -> int:
        return len(self.owner)

    @property
    def n_internal(self) -> int:
        return len(self.neighbour)

    def patch(self, name: str) -> tuple:
        for p in self.patches:
            if p[0] == name:
                return p
        raise KeyError(name)


def structured_hex(x, y, z, active=None, classify=None, names=None) -> PolyMeshArrays:
    """Hexahedra of the grid with node coordinates x, y, z (increasing), the
    cells where active[i, j, k] is true (all by default). Cells are numbered
    with i fastest, then j, then k. Boundary faces go to patches: `classify`
    maps (axis, side, cx, cy, cz) arrays of the boundary faces (axis 0/1/2,
    side -1 at a cell's lower face, +1 at its upper face, face centres) to
    indices into `names`; by default the six box patches xmin, xmax, ymin,
    ymax, zmin, zmax. Empty patches are dropped."""
    x, y, z = (np.asarray(v, np.float64) for v in (x, y, z))
    nx, ny, nz = len(x) - 1, len(y) - 1, len(z) - 1
    if min(nx, ny, nz) < 1 or any(np.any(np.diff(v) <= 0) for v in (x, y, z)):
        raise ValueError("structured_hex: node coordinates must increase, at least one cell per axis")
    act = np.ones((nx, ny, nz), bool) if active is None else np.asarray(active, bool)
    if act.shape != (nx, ny, nz):
        raise ValueError("structured_hex: active mask shape mismatch")
    if names is None:
        names = ["xmin", "xmax", "ymin", "ymax", "zmin", "zmax"]
        classify = lambda axis, side, cx, cy, cz: 2 * axis + (side > 0)  # noqa: E731

    # Cell ids, i fastest.
    flat = act.transpose(2, 1, 0).reshape(-1)
    cid = np.full(flat.shape, -1, np.int64)
    cid[flat] = np.arange(int(flat.sum()))
    cid = cid.reshape(nz, ny, nx).transpose(2, 1, 0)  # cid[i, j, k]
    n_cells = int(flat.sum())

    # Nodes used by an active cell, numbered i fastest.
    used = np.zeros((nx + 1, ny + 1, nz + 1), bool)
    for di in (0, 1):
        for dj in (0, 1):
            for dk in (0, 1):
                used[di:nx + di, dj:ny + dj, dk:nz + dk] |= act
    uflat = used.transpose(2, 1, 0).reshape(-1)
    nid = np.full(uflat.shape, -1, np.int64)
    nid[uflat] = np.arange(int(uflat.sum()))
    nid = nid.reshape(nz + 1, ny + 1, nx + 1).transpose(2, 1, 0)
    gi, gj, gk = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    sel = used.transpose(2, 1, 0).reshape(-1)
    pts = np.stack([x[gi.transpose(2, 1, 0).reshape(-1)[sel]], y[gj.transpose(2, 1, 0).reshape(-1)[sel]],
                    z[gk.transpose(2, 1, 0).reshape(-1)[sel]]], axis=1)

    internal_faces, boundary_faces = [], []
    coords = (x, y, z)
    for axis in range(3):
        n_a = (nx, ny, nz)[axis]
        # Move the face axis first: planes p = 0..n_a, the other two cell indices.
        a = np.moveaxis(act, axis, 0)
        c = np.moveaxis(cid, axis, 0)
        pad = np.zeros((1,) + a.shape[1:], bool)
        lower = np.concatenate([pad, a], 0)   # cell below plane p is active
        upper = np.concatenate([a, pad], 0)   # cell above plane p is active
        cpad = np.full((1,) + c.shape[1:], -1, np.int64)
        lower_id = np.concatenate([cpad, c], 0)
        upper_id = np.concatenate([c, cpad], 0)
        p, s, t = np.meshgrid(np.arange(n_a + 1), np.arange(a.shape[1]), np.arange(a.shape[2]), indexing="ij")
        # The quad's nodes in the winding whose area vector is +axis.
        if axis == 0:
            quad = [(p, s, t), (p, s + 1, t), (p, s + 1, t + 1), (p, s, t + 1)]
        elif axis == 1:  # s = i, t = k: z then x gives +y
            quad = [(s, p, t), (s, p, t + 1), (s + 1, p, t + 1), (s + 1, p, t)]
        else:  # s = i, t = j
            quad = [(s, t, p), (s + 1, t, p), (s + 1, t + 1, p), (s, t + 1, p)]
        nodes = np.stack([nid[q] for q in quad], -1)  # (planes, s, t, 4)
        both = lower & upper
        internal_faces.append((lower_id[both], upper_id[both], nodes[both]))
        for side, mask, owner_id in ((+1, lower & ~upper, lower_id), (-1, upper & ~lower, upper_id)):
            q = nodes[mask]
            if side < 0:
                q = q[:, ::-1]
            others = [ax for ax in range(3) if ax != axis]
            centre = [None, None, None]
            centre[axis] = coords[axis][p[mask]]
            centre[others[0]] = 0.5 * (coords[others[0]][s[mask]] + coords[others[0]][s[mask] + 1])
            centre[others[1]] = 0.5 * (coords[others[1]][t[mask]] + coords[others[1]][t[mask] + 1])
            patch = np.asarray(classify(axis, side, *centre), np.int64)
            patch = np.broadcast_to(patch, (int(mask.sum()),))
            boundary_faces.append((owner_id[mask], q, patch))

    own_i = np.concatenate([f[0] for f in internal_faces])
    nei_i = np.concatenate([f[1] for f in internal_faces])
    nod_i = np.concatenate([f[2] for f in internal_faces])
    own_b = np.concatenate([f[0] for f in boundary_faces])
    nod_b = np.concatenate([f[1] for f in boundary_faces])
    pat_b = np.concatenate([f[2] for f in boundary_faces])
    if np.any(pat_b < 0) or np.any(pat_b >= len(names)):
        raise ValueError("structured_hex: a boundary face has no patch")
    order = np.argsort(pat_b, kind="stable")
    own_b, nod_b, pat_b = own_b[order], nod_b[order], pat_b[order]
    patches, start = [], len(own_i)
    for k, name in enumerate(names):
        size = int(np.count_nonzero(pat_b == k))
        if size:
            patches.append((name, start, size))
            start += size
    owner = np.concatenate([own_i, own_b]).astype(np.int32)
    nodes = np.concatenate([nod_i, nod_b]).astype(np.int32)
    return PolyMeshArrays(points=np.ascontiguousarray(pts), face_offsets=np.arange(0, 4 * len(owner) + 1, 4,
                                                                                   dtype=np.int32),
                          face_nodes=np.ascontiguousarray(nodes.reshape(-1)), owner=owner,
                          neighbour=nei_i.astype(np.int32), n_cells=n_cells, patches=patches)


def box_hex(nx, ny, nz, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)) -> PolyMeshArrays:
    """A uniform hexahedral box with the six patches xmin ... zmax."""
    return structured_hex(*(np.linspace(lo[a], hi[a], n + 1) for a, n in enumerate((nx, ny, nz))))


def graded(length: float, first: float, growth: float, largest: float) -> np.ndarray:
    """Cell sizes from a wall: `first`, growing geometrically by `growth`
    until `largest`, then constant, scaled down uniformly so they sum to
    `length`."""
    sizes, s = [], first
    while sum(sizes) + s < length:
        sizes.append(s)
        s = min(s * growth, largest)
    sizes = np.asarray(sizes if sizes else [length])
    rest = length - sizes.sum()
    if rest > 0.5 * sizes[-1]:
        sizes = np.append(sizes, rest)
    return sizes * (length / sizes.sum())


def two_sided(length: float, first: float, growth: float, largest: float) -> np.ndarray:
    """Cell sizes graded from walls at both ends of a segment (symmetric)."""
    half = graded(0.5 * length, first, growth, largest)
    return np.concatenate([half, half[::-1]])


def nodes_from(start: float, sizes) -> np.ndarray:
    return start + np.concatenate([[0.0], np.cumsum(sizes)])


@dataclass
class StepGeometry:
    """Backward-facing step channel: step height h, the inlet `upstream`
    lengths of h before the step, the outlet `downstream` lengths of h after
    it, channel height above the step `channel` lengths of h (8h: the
    expansion ratio 1.125 of Driver and Seegmiller's step), span `span`
    lengths of h between symmetry planes."""
    h: float = 0.1
    upstream: float = 5.0
    downstream: float = 30.0
    channel: float = 8.0
    span: float = 4.0


def first_cell_height(speed: float, nu: float, x: float, y_plus: float = 1.0) -> float:
    """Wall-normal size of the first cell for a target y+ from the turbulent
    flat-plate skin friction Cf = 0.026 Re_x^(-1/7) at distance x."""
    re_x = speed * x / nu
    cf = 0.026 * re_x ** (-1.0 / 7.0)
    u_tau = speed * math.sqrt(0.5 * cf)
    return y_plus * nu / u_tau


def backward_step(geometry: StepGeometry = StepGeometry(), first: float = 1.7e-5, growth: float = 1.15,
                  largest: float = 0.02, nz: int = 80, outlet_largest: float | None = None) -> PolyMeshArrays:
    """Hexahedral backward-facing step: walls graded geometrically from
    `first` by `growth` up to `largest`: the upstream floor (y = 0, x < 0),
    the step face (x = 0, y < 0), the downstream floor (y = -h) and the top
    wall (y = channel h). Patches: inlet (x = -upstream h), outlet
    (x = downstream h), wall, and symmetry at z = 0 and z = span h."""
    g = geometry
    h = g.h
    big = outlet_largest or largest
    xu = graded(g.upstream * h, first, growth, largest)[::-1]  # fine at the step
    xd = graded(g.downstream * h, first, growth, big)
    x = nodes_from(-g.upstream * h, np.concatenate([xu, xd]))
    y_low = two_sided(h, first, growth, largest)
    y_high = two_sided(g.channel * h, first, growth, largest)
    y = nodes_from(-h, np.concatenate([y_low, y_high]))
    z = np.linspace(0.0, g.span * h, nz + 1)
    x[len(xu)] = 0.0
    y[len(y_low)] = 0.0
    xc = 0.5 * (x[:-1] + x[1:])
    yc = 0.5 * (y[:-1] + y[1:])
    active = ~((xc[:, None] < 0) & (yc[None, :] < 0))
    active = np.repeat(active[:, :, None], nz, axis=2)
    names = ["inlet", "outlet", "wall", "symmetry"]
    tol = 1e-9 * h

    def classify(axis, side, cx, cy, cz):
        out = np.full(np.shape(cx), 2, np.int64)  # walls by default
        if axis == 0:
            out[(side < 0) & (cx < -g.upstream * h + tol)] = 0
            out[(side > 0) & (cx > g.downstream * h - tol)] = 1
        if axis == 2:
            out[:] = 3
        return out

    return structured_hex(x, y, z, active, classify, names)
