"""Synthetic seeded meshes and states for the lid-driven cavity and channel
workloads (input generation; the assembly itself lives in the HIP library).

Element corner order follows the UG4 reference elements: triangles and quads
counter-clockwise from the lower-left corner, tetrahedra (0,0,0),(1,0,0),
(0,1,0),(0,0,1), hexahedra bottom face counter-clockwise then the top face,
prisms the bottom triangle counter-clockwise then the triangle above it.
Node numbering is lexicographic (x fastest); DoF = node*(dim+1)+component.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

TRI, QUAD, TET, HEX, PRISM = 0, 1, 2, 3, 4
ELEM_NAMES = {TRI: "triangle", QUAD: "quadrilateral", TET: "tetrahedron", HEX: "hexahedron", PRISM: "prism"}
DIM = {TRI: 2, QUAD: 2, TET: 3, HEX: 3, PRISM: 3}
NSH = {TRI: 3, QUAD: 4, TET: 4, HEX: 8, PRISM: 6}
# element sides in UG4 reference-element numbering (corners of each side)
SIDES = {
    TRI: [(0, 1), (1, 2), (2, 0)],
    QUAD: [(0, 1), (1, 2), (2, 3), (3, 0)],
    TET: [(0, 2, 1), (1, 2, 3), (0, 3, 2), (0, 1, 3)],
    HEX: [(0, 3, 2, 1), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (0, 4, 7, 3), (4, 5, 6, 7)],
    PRISM: [(0, 2, 1), (0, 1, 4, 3), (1, 2, 5, 4), (2, 0, 3, 5), (3, 4, 5)],
}


def boundary_sides(mesh: "Mesh", nodes) -> tuple:
    """(element, side) pairs of every element side whose corners all lie in
    `nodes` (e.g. the outflow plane).  Returns int32 arrays.  On a mixed
    mesh the elements are numbered group after group."""
    mask = np.zeros(mesh.nv, dtype=bool)
    mask[np.asarray(nodes, dtype=np.int64)] = True
    belem, bside = [], []
    groups = mesh.groups if getattr(mesh, "groups", None) is not None else [(mesh.et, mesh.conn)]
    e0 = 0
    for et, conn in groups:
        for s, corners in enumerate(SIDES[et]):
            on = np.all(mask[conn[:, list(corners)]], axis=1)
            e = np.nonzero(on)[0]
            belem.append(e + e0)
            bside.append(np.full(len(e), s))
        e0 += conn.shape[0]
    belem = np.concatenate(belem)
    bside = np.concatenate(bside)
    o = np.lexsort((bside, belem))
    return belem[o].astype(np.int32), bside[o].astype(np.int32)


@dataclass
class Mesh:
    et: int
    xyz: np.ndarray         # [nv, dim] float64
    conn: np.ndarray        # [ne, nsh] int32
    boundary: dict          # name -> node index array
    shape: tuple = ()       # structured cell counts (nx, ny[, nz])
    recipe: dict | None = None  # how to rebuild the mesh at another resolution

    @property
    def dim(self) -> int:
        return DIM[self.et]

    @property
    def nv(self) -> int:
        return self.xyz.shape[0]

    @property
    def ne(self) -> int:
        return self.conn.shape[0]


def _jitter(xyz: np.ndarray, interior: np.ndarray, h: float, amount: float, seed: int) -> np.ndarray:
    if amount <= 0:
        return xyz
    rng = np.random.default_rng(seed)
    d = rng.uniform(-amount * h, amount * h, size=xyz.shape)
    out = xyz.copy()
    out[interior] += d[interior]
    return out


def cavity_2d(n: int, et: int = QUAD, jitter: float = 0.0, seed: int = 0, lx: float = 1.0) -> Mesh:
    """Unit square, n x n cells; quads, or each square split into 2 triangles
    along the (i,j)-(i+1,j+1) diagonal.  Interior nodes jittered uniformly
    within +-jitter*h per coordinate."""
    h = lx / n
    i, j = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="xy")
    xyz = np.stack([i.ravel() * h, j.ravel() * h], axis=1).astype(np.float64)
    v = lambda a, b: b * (n + 1) + a
    ci, cj = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")
    ci, cj = ci.ravel(), cj.ravel()
    v00, v10, v11, v01 = v(ci, cj), v(ci + 1, cj), v(ci + 1, cj + 1), v(ci, cj + 1)
    if et == QUAD:
        conn = np.stack([v00, v10, v11, v01], axis=1)
    elif et == TRI:
        t1 = np.stack([v00, v10, v11], axis=1)
        t2 = np.stack([v00, v11, v01], axis=1)
        conn = np.stack([t1, t2], axis=1).reshape(-1, 3)
    else:
        raise ValueError("2D cavity needs TRI or QUAD")
    ii, jj = i.ravel(), j.ravel()
    interior = (ii > 0) & (ii < n) & (jj > 0) & (jj < n)
    xyz = _jitter(xyz, interior, h, jitter, seed)
    lid = np.nonzero(jj == n)[0]
    wall = np.nonzero(((ii == 0) | (ii == n) | (jj == 0)) & (jj < n))[0]
    return Mesh(et, xyz, conn.astype(np.int32), {"lid": lid, "wall": wall}, (n, n),
                {"fn": "cavity_2d", "et": et, "lx": lx})


def cavity_3d(n: int, et: int = HEX, jitter: float = 0.0, seed: int = 0) -> Mesh:
    """Unit cube, n^3 hexahedra (or 6 Kuhn tetrahedra per hexahedron)."""
    return box_3d((n, n, n), (1.0, 1.0, 1.0), et=et, jitter=jitter, seed=seed, kind="cavity")


# Kuhn split of the unit cube into 6 tetrahedra sharing the (0,0,0)-(1,1,1)
# diagonal; each listed with positive orientation.
_KUHN = [
    (0, 1, 2, 6), (0, 2, 3, 6), (0, 3, 7, 6), (0, 7, 4, 6), (0, 4, 5, 6), (0, 5, 1, 6),
]


# a hexahedron split along its bottom/top diagonal 0-2 / 4-6 into two
# prisms (triangle at z = 0 counter-clockwise from +z, then the one above)
_PRISMS = [(0, 1, 2, 4, 5, 6), (0, 2, 3, 4, 6, 7)]


def box_3d(cells, lengths, et: int = HEX, jitter: float = 0.0, seed: int = 0, kind: str = "cavity",
           keep_cell=None, solid=None) -> Mesh:
    """Box [0,lx]x[0,ly]x[0,lz] of nx*ny*nz hexahedra (or 6 Kuhn tetrahedra,
    or 2 prisms, each).  keep_cell(ci, cj, ck) or solid(x, y, z) (evaluated at cell
    centres; True = inside an obstacle) drop cells; their nodes stay in the
    numbering and are listed in boundary["unused"]."""
    nx, ny, nz = cells
    hx, hy, hz = lengths[0] / nx, lengths[1] / ny, lengths[2] / nz
    k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    ii, jj, kk = i.ravel(), j.ravel(), k.ravel()
    xyz = np.stack([ii * hx, jj * hy, kk * hz], axis=1).astype(np.float64)
    v = lambda a, b, c: (c * (ny + 1) + b) * (nx + 1) + a
    ck, cj, ci = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()
    if keep_cell is not None:
        m = keep_cell(ci, cj, ck)
        ci, cj, ck = ci[m], cj[m], ck[m]
    if solid is not None:
        m = ~solid((ci + 0.5) * hx, (cj + 0.5) * hy, (ck + 0.5) * hz)
        ci, cj, ck = ci[m], cj[m], ck[m]
    hexc = np.stack([v(ci, cj, ck), v(ci + 1, cj, ck), v(ci + 1, cj + 1, ck), v(ci, cj + 1, ck),
                     v(ci, cj, ck + 1), v(ci + 1, cj, ck + 1), v(ci + 1, cj + 1, ck + 1), v(ci, cj + 1, ck + 1)],
                    axis=1)
    if et == HEX:
        conn = hexc
    elif et == TET:
        conn = np.stack([hexc[:, list(t)] for t in _KUHN], axis=1).reshape(-1, 4)
    elif et == PRISM:
        conn = np.stack([hexc[:, list(t)] for t in _PRISMS], axis=1).reshape(-1, 6)
    else:
        raise ValueError("3D box needs TET, HEX or PRISM")
    interior = (ii > 0) & (ii < nx) & (jj > 0) & (jj < ny) & (kk > 0) & (kk < nz)
    xyz = _jitter(xyz, interior, min(hx, hy, hz), jitter, seed)
    if kind == "cavity":
        lid = np.nonzero(kk == nz)[0]
        wall = np.nonzero(((ii == 0) | (ii == nx) | (jj == 0) | (jj == ny) | (kk == 0)) & (kk < nz))[0]
        bnd = {"lid": lid, "wall": wall}
    else:
        inflow = np.nonzero(ii == 0)[0]
        outflow = np.nonzero(ii == nx)[0]
        wall = np.nonzero(((jj == 0) | (jj == ny) | (kk == 0) | (kk == nz)) & (ii > 0) & (ii < nx))[0]
        bnd = {"inflow": inflow, "outflow": outflow, "wall": wall}
    used = np.zeros(xyz.shape[0], dtype=bool)
    used[conn.ravel()] = True
    bnd["unused"] = np.nonzero(~used)[0]
    recipe = None if keep_cell is not None else {"fn": "box_3d", "et": et, "lengths": tuple(lengths), "kind": kind,
                                                 "solid": solid}
    return Mesh(et, xyz, conn.astype(np.int32), bnd, (nx, ny, nz), recipe)


@dataclass
class MixedMesh:
    """A mesh of several element types over one node set: groups of
    (element type, connectivity [ne, nsh]); elements are numbered group after
    group."""
    groups: list
    xyz: np.ndarray
    boundary: dict
    shape: tuple = ()

    @property
    def dim(self) -> int:
        return self.xyz.shape[1]

    @property
    def nv(self) -> int:
        return self.xyz.shape[0]

    @property
    def ne(self) -> int:
        return sum(c.shape[0] for _, c in self.groups)


def mixed_2d(n: int, jitter: float = 0.0, seed: int = 0, first: int = QUAD) -> MixedMesh:
    """The n x n cavity of cavity_2d with every third cell ((i + 2j) % 3 ==
    0) split into two triangles along its (i,j)-(i+1,j+1) diagonal, the
    others quadrilaterals; `first` names the group listed first."""
    m = cavity_2d(n, QUAD, jitter, seed)
    ci, cj = np.arange(m.ne) % n, np.arange(m.ne) // n
    split = (ci + 2 * cj) % 3 == 0
    quads = m.conn[~split]
    q = m.conn[split]
    tris = np.stack([q[:, [0, 1, 2]], q[:, [0, 2, 3]]], axis=1).reshape(-1, 3)
    groups = [(QUAD, quads), (TRI, tris.astype(np.int32))]
    if first == TRI:
        groups = groups[::-1]
    return MixedMesh(groups, m.xyz, m.boundary, (n, n))


def mixed_3d(n: int, kind: str = "hex_prism", jitter: float = 0.0, seed: int = 0) -> MixedMesh:
    """The n^3 cavity of cavity_3d with conforming element types:
    "hex_prism": the cells with i < n/2 split into two prisms each (their
      quadrilateral sides meet the hexahedra whole), the others hexahedra;
    "prism_tet": the cells with k < n/2 split into two prisms, the cells
      above into six Kuhn tetrahedra (both split the horizontal faces along
      the same diagonal).
    A hexahedron and a tetrahedron cannot share a face (a quadrilateral
    against two triangles): that needs a pyramid between them."""
    m = box_3d((n, n, n), (1.0, 1.0, 1.0), et=HEX, jitter=jitter, seed=seed)
    c = np.arange(m.ne)
    ci, ck = c % n, c // (n * n)
    low_i, low_k = ci < n // 2, ck < n // 2
    if kind == "hex_prism":
        sel = {HEX: ~low_i, PRISM: low_i}
        order = [HEX, PRISM]
    elif kind == "prism_tet":
        sel = {PRISM: low_k, TET: ~low_k}
        order = [PRISM, TET]
    else:
        raise ValueError(kind)
    groups = []
    for et in order:
        h = m.conn[sel[et]]
        if et == HEX:
            conn = h
        elif et == PRISM:
            conn = np.stack([h[:, list(t)] for t in _PRISMS], axis=1).reshape(-1, 6)
        else:
            conn = np.stack([h[:, list(t)] for t in _KUHN], axis=1).reshape(-1, 4)
        groups.append((et, conn.astype(np.int32)))
    return MixedMesh(groups, m.xyz, m.boundary, (n, n, n))


def split_groups(mesh: Mesh, cut: int) -> MixedMesh:
    """One element type as two groups: elements [0, cut) and [cut, ne)."""
    return MixedMesh([(mesh.et, mesh.conn[:cut]), (mesh.et, mesh.conn[cut:])], mesh.xyz, mesh.boundary, mesh.shape)


def _grid_index(shape):
    """Structured node coordinates (i, j[, k]) of a lexicographic numbering."""
    dims = [s + 1 for s in shape]
    idx = np.indices(dims[::-1]).reshape(len(dims), -1)[::-1]
    return idx  # idx[0] = i (fastest)


def coarsen(mesh: Mesh):
    """Next coarser level of a structured mesh (every cell count halved).

    Returns (coarse, prolongation, injection): the coarse mesh (same element
    type and boundary sets; node positions taken from the fine mesh at the
    coinciding nodes, so jitter carries over), the node-level prolongation as
    CSR (rowptr, colidx, weights) with rows = fine nodes and columns = coarse
    nodes (multilinear interpolation in grid index space), and the injection
    coarse node -> fine node."""
    if mesh.recipe is None or any(c % 2 for c in mesh.shape):
        raise ValueError("mesh cannot be coarsened (needs a recipe and even cell counts)")
    r = mesh.recipe
    cshape = tuple(c // 2 for c in mesh.shape)
    if r["fn"] == "cavity_2d":
        coarse = cavity_2d(cshape[0], r["et"], lx=r["lx"])
    else:
        coarse = box_3d(cshape, r["lengths"], et=r["et"], kind=r["kind"], solid=r["solid"])
        if "boxes" in r:
            coarse.recipe["boxes"] = r["boxes"]
            coarse.boundary["obstacle"] = obstacle_surface_nodes(coarse)
    fdims = [c + 1 for c in mesh.shape]
    cdims = [c + 1 for c in cshape]
    fidx = _grid_index(mesh.shape)
    cidx = _grid_index(cshape)
    # injection: coarse (I, J, K) sits on fine (2I, 2J, 2K)
    inj = np.zeros(coarse.nv, dtype=np.int64)
    stride = 1
    for d in range(len(fdims)):
        inj += 2 * cidx[d] * stride
        stride *= fdims[d]
    coarse.xyz = mesh.xyz[inj].copy()
    # prolongation: tensor product of 1D weights
    nf = mesh.nv
    rows = [np.arange(nf)]
    cols = [np.zeros(nf, dtype=np.int64)]
    wts = [np.ones(nf)]
    cstride = 1
    for d in range(len(fdims)):
        i = fidx[d]
        lo = i // 2
        odd = (i % 2) == 1
        nr, nc_, nw = [], [], []
        for rr, cc, ww in zip(rows, cols, wts):
            # even: one parent weight 1; odd: parents lo and lo+1 weight 1/2
            nr.append(rr); nc_.append(cc + lo[rr] * cstride); nw.append(ww * np.where(odd[rr], 0.5, 1.0))
            m = odd[rr]
            nr.append(rr[m]); nc_.append(cc[m] + (lo[rr[m]] + 1) * cstride); nw.append(ww[m] * 0.5)
        rows = [np.concatenate(nr)]
        cols = [np.concatenate(nc_)]
        wts = [np.concatenate(nw)]
        cstride *= cdims[d]
    rows, cols, wts = rows[0], cols[0], wts[0]
    order = np.lexsort((cols, rows))
    rows, cols, wts = rows[order], cols[order], wts[order]
    rowptr = np.zeros(nf + 1, dtype=np.int64)
    np.add.at(rowptr, rows + 1, 1)
    rowptr = np.cumsum(rowptr)
    return coarse, (rowptr, cols.astype(np.int32), wts), inj


def obstacle_boxes(n_boxes: int, seed: int, lengths=(2.0, 1.0, 1.0), grid: float = 1.0 / 6.0):
    """Seeded axis-aligned obstacle boxes inside a channel, with faces on a
    grid of spacing `grid` (so every multigrid level resolves them) and a
    free inflow and outflow section.  Returns a list of (lo, hi) corners."""
    rng = np.random.default_rng(seed)
    boxes = []
    L = np.asarray(lengths)
    for _ in range(n_boxes):
        size = rng.integers(1, 3, size=3) * grid
        lo = np.array([rng.integers(2, int(round((L[0] - 0.5) / grid)) - 1) * grid,
                       rng.integers(0, int(round((L[1] - size[1]) / grid)) + 1) * grid,
                       rng.integers(0, int(round((L[2] - size[2]) / grid)) + 1) * grid])
        boxes.append((lo, lo + size))
    return boxes


def solid_from_boxes(boxes):
    def solid(x, y, z):
        m = np.zeros(np.shape(x), dtype=bool)
        for lo, hi in boxes:
            m |= (x > lo[0]) & (x < hi[0]) & (y > lo[1]) & (y < hi[1]) & (z > lo[2]) & (z < hi[2])
        return m
    return solid


def channel_3d(cells=(96, 48, 48), lengths=(2.0, 1.0, 1.0), et: int = TET, n_boxes: int = 2, seed: int = 0) -> Mesh:
    """Channel [0,2]x[0,1]x[0,1] of hexahedra split into 6 Kuhn tetrahedra
    each, with seeded obstacle boxes removed; inflow x = 0, outflow x = 2."""
    boxes = obstacle_boxes(n_boxes, seed, lengths)
    m = box_3d(cells, lengths, et=et, kind="channel", solid=solid_from_boxes(boxes))
    m.boundary["obstacle"] = obstacle_surface_nodes(m)
    m.recipe["boxes"] = boxes
    return m


def obstacle_surface_nodes(mesh: Mesh) -> np.ndarray:
    """Used nodes with a removed (solid) cell among their incident grid
    cells: the no-slip faces of the obstacles."""
    nx, ny, nz = mesh.shape
    lx, ly, lz = mesh.recipe["lengths"]
    solid = mesh.recipe["solid"]
    ck, cj, ci = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    removed = solid((ci + 0.5) * lx / nx, (cj + 0.5) * ly / ny, (ck + 0.5) * lz / nz)  # [nz, ny, nx]
    pad = np.zeros((nz + 2, ny + 2, nx + 2), dtype=bool)
    pad[1:-1, 1:-1, 1:-1] = removed
    touch = np.zeros((nz + 1, ny + 1, nx + 1), dtype=bool)
    for dk in (0, 1):
        for dj in (0, 1):
            for di in (0, 1):
                touch |= pad[dk:dk + nz + 1, dj:dj + ny + 1, di:di + nx + 1]
    used = np.zeros(mesh.nv, dtype=bool)
    used[mesh.conn.ravel()] = True
    return np.nonzero(touch.ravel() & used)[0]


def channel_dirichlet(mesh: Mesh, peak: float = 1.0):
    """Parabolic inflow and outflow u = 16 peak y(1-y) z(1-z) (fully
    developed, peak at the centre), no-slip walls and obstacle faces,
    every DoF of unused nodes fixed, pressure pinned at the outflow centre
    node.  Returns (dofs, vals), sorted by DoF."""
    nc = 4
    x, y, z = mesh.xyz[:, 0], mesh.xyz[:, 1], mesh.xyz[:, 2]
    prof = 16.0 * peak * y * (1 - y) * z * (1 - z)
    vel_nodes = np.unique(np.concatenate([mesh.boundary["inflow"], mesh.boundary["outflow"], mesh.boundary["wall"],
                                          mesh.boundary.get("obstacle", np.zeros(0, np.int64))]))
    dofs = (vel_nodes[:, None] * nc + np.arange(3)[None, :]).ravel()
    vals = np.zeros((len(vel_nodes), 3))
    io = np.isin(vel_nodes, np.concatenate([mesh.boundary["inflow"], mesh.boundary["outflow"]]))
    wall = np.isin(vel_nodes, np.concatenate([mesh.boundary["wall"], mesh.boundary.get("obstacle", np.zeros(0, np.int64))]))
    vals[io & ~wall, 0] = prof[vel_nodes[io & ~wall]]
    vals = vals.ravel()
    out = mesh.boundary["outflow"]
    centre = out[np.argmin((y[out] - 0.5) ** 2 + (z[out] - 0.5) ** 2)]
    dofs = np.append(dofs, centre * nc + 3)
    vals = np.append(vals, 0.0)
    unused = mesh.boundary.get("unused", np.zeros(0, np.int64))
    if len(unused):
        ex = (np.asarray(unused)[:, None] * nc + np.arange(nc)[None, :]).ravel()
        dofs = np.append(dofs, ex)
        vals = np.append(vals, np.zeros(len(ex)))
    o = np.argsort(dofs)
    return dofs[o].astype(np.int64), vals[o]


def channel_initial_state(mesh: Mesh, dirichlet, peak: float = 1.0) -> np.ndarray:
    """Parabolic profile in the whole channel, zero in and on obstacles,
    Dirichlet values imposed; pressure zero."""
    y, z = mesh.xyz[:, 1], mesh.xyz[:, 2]
    u = np.zeros((mesh.nv, 4))
    u[:, 0] = 16.0 * peak * y * (1 - y) * z * (1 - z)
    dofs, vals = dirichlet
    u.reshape(-1)[dofs] = vals
    return u


def random_state(mesh: Mesh, seed: int = 1, scale: float = 1.0) -> np.ndarray:
    """Seeded random nodal state [nv, dim+1] (velocity, pressure)."""
    rng = np.random.default_rng(seed)
    return scale * rng.uniform(-1.0, 1.0, size=(mesh.nv, mesh.dim + 1))


def divergence_free_perturbation(mesh: Mesh, amplitude: float = 1e-3, seed: int = 0, modes: int = 3) -> np.ndarray:
    """Velocity = curl of a seeded smooth vector potential that vanishes on the
    unit-cube boundary; pressure zero.  Returns [nv, 4]."""
    rng = np.random.default_rng(seed)
    x, y, z = mesh.xyz[:, 0], mesh.xyz[:, 1], mesh.xyz[:, 2]
    u = np.zeros((mesh.nv, 4))
    for _ in range(modes):
        a = rng.uniform(-1, 1, size=3)
        kx, ky, kz = rng.integers(1, 4, size=3)
        # psi_c = a_c * B(x,y,z) with B = sin^2 products (zero value and gradient on the boundary)
        sx, sy, sz = np.sin(np.pi * kx * x), np.sin(np.pi * ky * y), np.sin(np.pi * kz * z)
        dsx = 2 * np.pi * kx * sx * np.cos(np.pi * kx * x)
        dsy = 2 * np.pi * ky * sy * np.cos(np.pi * ky * y)
        dsz = 2 * np.pi * kz * sz * np.cos(np.pi * kz * z)
        B = sx**2 * sy**2 * sz**2
        Bx, By, Bz = dsx * sy**2 * sz**2, sx**2 * dsy * sz**2, sx**2 * sy**2 * dsz
        # curl(a B) = grad B x a
        gx, gy, gz = Bx, By, Bz
        u[:, 0] += gy * a[2] - gz * a[1]
        u[:, 1] += gz * a[0] - gx * a[2]
        u[:, 2] += gx * a[1] - gy * a[0]
    m = np.abs(u[:, :3]).max()
    if m > 0:
        u[:, :3] *= amplitude / m
    return u


def cavity_lid_state(mesh: Mesh, lid_speed: float = 1.0, base: np.ndarray | None = None) -> np.ndarray:
    """Apply the regularised lid (u = lid_speed along x on lid nodes except the
    lid's boundary edges, where u = 0) and no-slip walls to a state."""
    u = np.zeros((mesh.nv, mesh.dim + 1)) if base is None else base.copy()
    u[mesh.boundary["wall"], : mesh.dim] = 0.0
    lid = mesh.boundary["lid"]
    u[lid, : mesh.dim] = 0.0
    x = mesh.xyz[lid]
    inner = np.ones(len(lid), dtype=bool)
    for c in range(mesh.dim - 1):
        inner &= (x[:, c] > 1e-12) & (x[:, c] < 1 - 1e-12)
    u[lid[inner], 0] = lid_speed
    return u
