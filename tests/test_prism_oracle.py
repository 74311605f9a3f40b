"""Pins the oracle's prism (UG4 numbering; FV1Geometry restated for an
element whose corners all have three edges) by facts the geometry fixes:
- a right prism (a triangle translated along any vector) is affine: each of
  its six SCVs holds exactly one sixth of |base| x height, and the local ips
  map onto the global ones;
- every SCV is a closed surface: the signed SCVF normals of a corner plus
  the vector areas of its pieces of the element sides (computed here from
  the corner, edge midpoints and side centres) sum to zero, on jittered
  prisms too (this fixes the edge/side tables and the normal orientation);
- shapes are a partition of unity at every ip, and the global gradients
  reproduce every linear field exactly;
- the side/ray intersection lies on the cut side and behind the ip;
- LPS shapes reproduce a constant (they sum to one at a cut), the skewed
  upwind picks a corner of the cut side;
- an assembled prism mesh conserves: the A-part defect sums to zero over
  all nodes, component by component."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from plugin_navierstokes_amd import meshes as M

SIDES = M.SIDES[M.PRISM]
EDGES = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 4), (2, 5), (3, 4), (4, 5), (5, 3)]


def _right_prism(seed):
    rng = np.random.default_rng(seed)
    tri = rng.normal(size=(3, 3))
    t = rng.normal(size=3)
    return np.concatenate([tri, tri + t]), tri, t


def _jittered_prism(seed):
    rng = np.random.default_rng(seed)
    co = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1]], dtype=float)
    return co + rng.uniform(-0.12, 0.12, size=co.shape)


@pytest.mark.parametrize("seed", range(5))
def test_right_prism_volumes_and_ips(seed):
    co, tri, t = _right_prism(seed)
    G = O.geometry(O.PRISM, co)
    V = abs(np.dot(np.cross(tri[1] - tri[0], tri[2] - tri[0]), t)) / 2
    np.testing.assert_allclose(G["vol"], V / 6, rtol=1e-12)
    mapped = G["phi"] @ co
    np.testing.assert_allclose(mapped, G["ipg"], rtol=0, atol=1e-13 * np.abs(co).max())


def _side_piece_area(co, s, k):
    """Vector area of corner k's piece of side s: the quad [x_k, m(k, next),
    side centre, m(prev, k)], oriented away from the element centre."""
    cs = SIDES[s]
    j = cs.index(k)
    nxt, prv = cs[(j + 1) % len(cs)], cs[(j - 1) % len(cs)]
    fc = co[list(cs)].mean(0)
    a = 0.5 * np.cross(fc - co[k], 0.5 * (co[k] + co[nxt]) - 0.5 * (co[k] + co[prv]))
    if np.dot(a, fc - co.mean(0)) < 0:
        a = -a
    return a


@pytest.mark.parametrize("seed", range(6))
def test_scv_surfaces_are_closed(seed):
    co = _jittered_prism(seed) if seed else _right_prism(7)[0]
    G = O.geometry(O.PRISM, co)
    assert list(zip(G["from"], G["to"])) == EDGES
    scale = np.abs(G["n"]).max()
    for k in range(6):
        tot = np.zeros(3)
        for ip in range(9):
            if G["from"][ip] == k:
                tot += G["n"][ip]
            elif G["to"][ip] == k:
                tot -= G["n"][ip]
        for s in range(5):
            if k in SIDES[s]:
                tot += _side_piece_area(co, s, k)
        assert np.abs(tot).max() <= 1e-13 * scale, (k, tot)
    assert np.all(G["vol"] > 0)


@pytest.mark.parametrize("seed", range(4))
def test_shapes_and_gradients(seed):
    co = _jittered_prism(seed + 10)
    G = O.geometry(O.PRISM, co)
    np.testing.assert_allclose(G["phi"].sum(1), 1.0, rtol=1e-14)
    rng = np.random.default_rng(seed)
    a, b = rng.normal(size=3), rng.normal()
    f = co @ a + b
    for ip in range(9):
        np.testing.assert_allclose(G["grad"][ip].T @ f, a, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("seed", range(4))
def test_side_ray_and_upwinds(seed):
    co = _jittered_prism(seed + 20)
    G = O.geometry(O.PRISM, co)
    rng = np.random.default_rng(seed)
    for ip in range(9):
        v = rng.normal(size=3)
        side, gi, li = O.side_ray(O.PRISM, co, G["ipg"][ip], v)
        cs = list(SIDES[side])
        # behind the ip (upwind) and on the cut side's (bilinear) surface
        assert np.dot(gi - G["ipg"][ip], v) < 0
        phi = np.array([(1 - li[0] - li[1], li[0], li[1])[k % 3] * ((1 - li[2]) if k < 3 else li[2]) for k in range(6)])
        np.testing.assert_allclose(phi[[k for k in range(6) if k not in cs]], 0.0, atol=1e-12)
        assert np.all(phi[cs] >= -1e-9)
        # on a triangular side the cut point is the prism map of its local
        # coordinates (a quadrilateral side is cut on its two triangles)
        assert np.linalg.norm(phi @ co - gi) <= 1e-9 or len(cs) == 4
    vel = np.tile(rng.normal(size=3), (9, 1))
    U, L = O.upwind("lps", O.PRISM, co, vel)
    np.testing.assert_allclose(U.sum(1), 1.0, rtol=1e-12)
    U, L = O.upwind("skewed", O.PRISM, co, vel)
    assert np.all((U == 0) | (U == 1)) and np.all(U.sum(1) == 1)


@pytest.mark.parametrize("upw", ["full", "lps", "skewed"])
def test_assembled_mesh_conserves(upw):
    mesh = M.cavity_3d(2, M.PRISM, 0.1, seed=3)
    assert mesh.conn.shape == (16, 6)
    u = M.random_state(mesh, 4)
    p = O.Params(dim=3, upwind=upw, stab="fields", nu=0.02)
    J, d = O.assemble(p, O.PRISM, mesh.xyz, mesh.conn, u)
    dd = d.reshape(-1, 4)
    scale = np.abs(dd).sum(0)
    assert np.all(np.abs(dd.sum(0)) <= 1e-12 * scale)
    assert np.all(np.isfinite(J))
