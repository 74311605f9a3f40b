"""Pins the oracle's mixed-element assembly (several element types over one
node set, as ugcore's element loop runs the disc once per type of a
subset, navier_stokes_fv1.cpp:136-182 and :1515-1551):
- the mesh generators are conforming: every interior face is shared by
  exactly two elements with the same corner set (a quadrilateral never
  meets two triangles);
- the union pattern is, row by row, the union of the single-type patterns,
  and equals the single-type pattern for one group;
- the A-part defect conserves over the whole mixed mesh;
- with zero velocity and a linear pressure p = b.x + c the momentum rows
  of every interior node are b times the node's control volume (the SCVs of
  all its elements, of every type): the SCV surfaces close across the
  interfaces between element types;
- splitting one element type into two groups changes nothing."""
from __future__ import annotations

from collections import Counter

import numpy as np
import pytest

import oracle as O
from plugin_navierstokes_amd import meshes as M

MIXED = {
    "quad_tri": lambda: M.mixed_2d(6, 0.0, 1),
    "tri_quad": lambda: M.mixed_2d(5, 0.0, 2, first=M.TRI),
    "hex_prism": lambda: M.mixed_3d(4, "hex_prism", 0.0, 1),
    "prism_tet": lambda: M.mixed_3d(4, "prism_tet", 0.0, 2),
}
JITTERED = {
    "quad_tri": lambda: M.mixed_2d(6, 0.15, 1),
    "hex_prism": lambda: M.mixed_3d(4, "hex_prism", 0.1, 1),
    "prism_tet": lambda: M.mixed_3d(4, "prism_tet", 0.1, 2),
}


def _faces(mm):
    """Every element side as a sorted corner tuple, with its element type."""
    out = []
    for et, conn in mm.groups:
        for s in M.SIDES[et]:
            for row in conn[:, list(s)]:
                out.append(tuple(sorted(int(x) for x in row)))
    return out


@pytest.mark.parametrize("name", sorted(MIXED))
def test_meshes_are_conforming(name):
    mm = MIXED[name]()
    assert len(mm.groups) == 2 and mm.groups[0][0] != mm.groups[1][0]
    cnt = Counter(_faces(mm))
    assert max(cnt.values()) <= 2
    # a face seen once lies on the box boundary
    bnd = set(np.concatenate([mm.boundary["lid"], mm.boundary["wall"]]).tolist())
    for f, c in cnt.items():
        if c == 1:
            assert all(v in bnd for v in f), f
    # no side of one element lies inside a larger side of another
    # (quadrilateral against triangles): each triangle's corners are not a
    # subset of a quadrilateral's
    if mm.dim == 3:
        quads = {f for f in cnt if len(f) == 4}
        tris = [f for f in cnt if len(f) == 3]
        for t in tris:
            assert not any(set(t) <= set(q) for q in quads), t
    # every node is used and the volumes add up to the box
    used = np.zeros(mm.nv, bool)
    vol = 0.0
    for et, conn in mm.groups:
        used[conn.ravel()] = True
        vol += sum(O.geometry(et, mm.xyz[c])["vol"].sum() for c in conn)
    assert used.all()
    assert abs(vol - 1.0) < 1e-12


@pytest.mark.parametrize("name", sorted(MIXED))
def test_union_pattern(name):
    mm = MIXED[name]()
    rp, ci = O.pattern_mixed(mm.nv, mm.groups)
    nc = mm.dim + 1
    parts = [O.pattern(et, mm.nv, conn) for et, conn in mm.groups]
    for v in range(mm.nv):
        row = set(ci[rp[v * nc]:rp[v * nc + 1]].tolist())
        want = {v * nc + b for b in range(nc)}
        for prp, pci in parts:
            if prp[v * nc + 1] > prp[v * nc]:
                want |= set(pci[prp[v * nc]:prp[v * nc + 1]].tolist())
        assert row == want, v
    m = M.cavity_3d(3, M.PRISM, 0.1, 1)
    a, b = O.pattern(M.PRISM, m.nv, m.conn), O.pattern_mixed(m.nv, [(M.PRISM, m.conn)])
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("upw", ["full", "skewed", "lps"])
@pytest.mark.parametrize("name", sorted(JITTERED))
def test_mixed_defect_conserves(name, upw):
    mm = JITTERED[name]()
    rp, ci = O.pattern_mixed(mm.nv, mm.groups)
    u = M.random_state(mm, 4)
    p = O.Params(dim=mm.dim, upwind=upw, stab="fields", nu=0.02)
    J, d = O.assemble_mixed(p, mm.groups, mm.xyz, u, rp, ci)
    dd = d.reshape(-1, mm.dim + 1)
    assert np.all(np.abs(dd.sum(0)) <= 1e-12 * np.abs(dd).sum(0))
    assert np.all(np.isfinite(J))


@pytest.mark.parametrize("name", sorted(MIXED))
def test_linear_pressure_gives_control_volumes(name):
    mm = MIXED[name]()
    dim, nc = mm.dim, mm.dim + 1
    rng = np.random.default_rng(11)
    b, c = rng.normal(size=dim), rng.normal()
    u = np.zeros((mm.nv, nc))
    u[:, dim] = mm.xyz @ b + c
    rp, ci = O.pattern_mixed(mm.nv, mm.groups)
    p = O.Params(dim=dim, upwind="full", stab="fields", nu=0.02)
    _, d = O.assemble_mixed(p, mm.groups, mm.xyz, u, rp, ci)
    d = d.reshape(-1, nc)
    cv = np.zeros(mm.nv)
    for et, conn in mm.groups:
        for e in conn:
            cv[e] += O.geometry(et, mm.xyz[e])["vol"]
    bnd = np.zeros(mm.nv, bool)
    bnd[mm.boundary["lid"]] = bnd[mm.boundary["wall"]] = True
    inner = ~bnd
    assert inner.sum() > 0
    # every interface node between the two types is an inner node somewhere
    touch = [np.zeros(mm.nv, bool) for _ in mm.groups]
    for t, (_, conn) in zip(touch, mm.groups):
        t[conn.ravel()] = True
    assert np.any(touch[0] & touch[1] & inner)
    want = cv[inner, None] * b[None, :]
    np.testing.assert_allclose(d[inner, :dim], want, rtol=1e-12, atol=1e-13 * np.abs(want).max())


def test_one_type_in_two_groups_changes_nothing():
    m = M.cavity_3d(3, M.TET, 0.1, 4)
    mm = M.split_groups(m, m.ne // 3)
    u = M.random_state(m, 5)
    uold = M.random_state(m, 6)
    p = O.Params(dim=3, upwind="lps", stab="fields", nu=0.02)
    rp, ci = O.pattern(M.TET, m.nv, m.conn)
    J1, d1 = O.assemble(p, M.TET, m.xyz, m.conn, u, uold, 0.01, s_m=1.0, parts=3, rowptr=rp, colidx=ci)
    J2, d2 = O.assemble_mixed(p, mm.groups, m.xyz, u, rp, ci, uold=uold, dt=0.01, s_m=1.0, parts=3)
    Ja, da = O.assemble_abs(p, M.TET, m.xyz, m.conn, u, uold, 0.01, s_m=1.0, parts=3, rowptr=rp, colidx=ci)
    assert np.all(np.abs(J1 - J2) <= 1e-14 * Ja) and np.all(np.abs(d1 - d2) <= 1e-14 * da)
