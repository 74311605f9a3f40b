"""Pins the oracle's regular upwind (upwind.cpp:792-964, 2D).

Its ray/SCV intersection is ugcore's SCVFofSCVRayIntersection, which the
reference tree does not hold, so the side numbering and the direction of
lambda on each side are a stated convention (oracle/fv1_oracle.c, before
regular_upwind); parity with ugcore itself stays unpinned.  What the tests
fix independently of the restated formulas:
- the upwind point sum_k U_k x_k + sum_j W_j ip_j equals the geometric exit
  point of the ray ip - t v (t > 0) from the upwind SCV, computed here by a
  brute-force search over the SCV's sides, wherever the reference's
  branches place it there (an element side; on quadrilaterals also the
  other SCVF between its edge midpoint and its ip);
- the shapes sum to one (a constant velocity is reproduced) whenever the
  velocity is nonzero;
- hand-derived unit-square values for the branch between the ip and the
  element centre, for flow along an SCVF, and for zero velocity;
- 3D elements are refused with the reference's message."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O

UNIT = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])


def _geo(et, co):
    """SCVF ips, normals, edge midpoints and centre, built from the box
    construction (independent of the oracle's geometry code)."""
    n = len(co)
    c = co.mean(axis=0)
    m = np.array([0.5 * (co[k] + co[(k + 1) % n]) for k in range(n)])
    ip = 0.5 * (m + c)
    nrm = np.array([[(c - m[k])[1], -(c - m[k])[0]] for k in range(n)])
    for k in range(n):
        if nrm[k] @ (co[(k + 1) % n] - co[k]) < 0:
            nrm[k] = -nrm[k]
    return ip, nrm, m, c


def _exit_point(o, v, segs):
    """First crossing of the ray o - t v (t > 0) with the segments, by
    solving each 2x2 system; returns (segment index, point, lambda)."""
    best = None
    for s, (a, b) in enumerate(segs):
        A = np.array([b - a, v]).T
        if abs(np.linalg.det(A)) < 1e-14:
            continue
        lam, t = np.linalg.solve(A, o - a)
        if t > 1e-12 and -1e-9 <= lam <= 1 + 1e-9 and (best is None or t < best[3]):
            best = (s, a + lam * (b - a), lam, t)
    return best


def _upwind_point(U, W, co, ip, j):
    return U[j] @ co + W[j] @ ip


@pytest.mark.parametrize("et,seed", [(O.QUAD, 0), (O.QUAD, 1), (O.TRI, 2), (O.TRI, 3)])
def test_upwind_point_is_the_exit_point(et, seed):
    rng = np.random.default_rng(seed)
    checked = 0
    for trial in range(200):
        if et == O.QUAD:
            co = UNIT + rng.uniform(-0.15, 0.15, size=(4, 2))
        else:
            co = np.array([[0.0, 0.0], [1.0, 0.1], [0.2, 0.9]]) + rng.uniform(-0.1, 0.1, size=(3, 2))
        n = len(co)
        v = rng.normal(size=(n, 2))
        U, W, L = O.upwind("regular", et, co, v, ip_shapes=True)
        ip, nrm, m, c = _geo(et, co)
        for j in range(n):
            assert abs(U[j].sum() + W[j].sum() - 1.0) < 1e-12
            flux = v[j] @ nrm[j]
            k = j if flux > 0 else (j + 1) % n
            oth = (j - 1) % n if flux > 0 else (j + 1) % n
            # the SCV of corner k without the SCVF the ray starts on
            segs = [(m[oth], c), (co[k], m[k]), (m[(k - 1) % n], co[k])]
            hit = _exit_point(ip[j], v[j], segs)
            assert hit is not None
            s, x, lam, _ = hit
            up = _upwind_point(U, W, co, ip, j)
            if s == 0 and (lam > 0.5 or et == O.TRI):
                # between the other SCVF's ip and the centre the reference
                # blends the corners of SCVF j with that ip instead; and its
                # extrapolation through the opposite ip is exact only where
                # the centre halves the opposite edge midpoints (every
                # quadrilateral), not on triangles
                continue
            assert np.allclose(up, x, atol=1e-12), (trial, j, s, lam, up, x)
            assert abs(L[j] - np.linalg.norm(ip[j] - x)) < 1e-12
            checked += 1
    assert checked > 200


def test_unit_square_hand_values():
    # v = (0.5, -1) everywhere.  SCVF 0 (edge 0-1, ip (0.5, 0.25), normal
    # (0.5, 0)): flux 0.25 > 0, upwind SCV 0; the ray (0.5 - 0.5t, 0.25 + t)
    # crosses SCVF 3 ([m30 = (0, 0.5), c]) at (0.375, 0.5): lambda = 0.75 from
    # m30, past SCVF 3's ip (0.25, 0.5): U0 = U1 = 0.5 * 2 * 0.25 = 0.25,
    # W3 = 1 - 2 * 0.25 = 0.5.
    v = np.tile([0.5, -1.0], (4, 1))
    U, W, L = O.upwind("regular", O.QUAD, UNIT, v, ip_shapes=True)
    assert np.allclose(U[0], [0.25, 0.25, 0, 0]) and np.allclose(W[0], [0, 0, 0, 0.5])
    up = 0.25 * UNIT[0] + 0.25 * UNIT[1] + 0.5 * np.array([0.25, 0.5])
    assert abs(L[0] - np.linalg.norm(np.array([0.5, 0.25]) - up)) < 1e-15
    # SCVF 3 (edge 3-0, ip (0.25, 0.5), normal (0, -0.5)): flux 0.5 > 0,
    # upwind SCV 3 ([x3, m30, c, m23]); the ray (0.25 - 0.5t, 0.5 + t) leaves
    # at t = 0.5 through the corner x3 = (0, 1), the start of side 0 (lambda
    # 0: U3 = 1 - 0, U0 = 0) and the end of side 3 (lambda 1 from m23:
    # U2 = 1 - 0.5 * 2 = 0, U3 = 1); both give the corner x3.
    assert np.allclose(U[3], [0, 0, 0, 1.0]) and np.allclose(W[3], 0)
    # flow along SCVF 0 (v parallel to its normal's perpendicular): v = (0, 1)
    # gives the cross term 0 * 0.5 - 1 * 0.5 < 0, the mean of corners 0 and 1;
    # v = (0, -1) the mean of the neighbouring ips 3 and 1
    U, W, _ = O.upwind("regular", O.QUAD, UNIT, np.tile([0.0, 1.0], (4, 1)), ip_shapes=True)
    assert np.allclose(U[0], [0.5, 0.5, 0, 0]) and np.allclose(W[0], 0)
    U, W, _ = O.upwind("regular", O.QUAD, UNIT, np.tile([0.0, -1.0], (4, 1)), ip_shapes=True)
    assert np.allclose(U[0], 0) and np.allclose(W[0], [0, 0.5, 0, 0.5])
    # zero velocity: no upwind shapes; the length is then |ip| (the upwind
    # point is the origin), as in the reference
    U, W, L = O.upwind("regular", O.QUAD, UNIT + 2.0, np.zeros((4, 2)), ip_shapes=True)
    assert np.allclose(U, 0) and np.allclose(W, 0)
    assert abs(L[0] - np.linalg.norm([2.5, 2.25])) < 1e-15


def test_regular_upwind_is_2d_only():
    co = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    with pytest.raises(RuntimeError, match="only implemented for 2d"):
        O.upwind("regular", O.TET, co, np.ones((6, 3)), ip_shapes=True)
