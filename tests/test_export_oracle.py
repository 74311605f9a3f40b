"""Pins the oracle's data exports (navier_stokes_fv1.cpp:894-1491):
linear fields are reproduced exactly by the P1/Q1 and prism interpolations
(also on distorted hexahedra and prisms, where x(xi) is not affine), so values and global
gradients at SCVF ips, corners and local points have closed forms; the
"general case" velocity gradient is the reference-element gradient
(:1072-1080), A B on an affine element x = B xi + x0; the stabilised ip
velocity closes the continuity rows of the element defect, and without
stabilisation it is the interpolated velocity."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O

ETS = [O.TRI, O.QUAD, O.TET, O.HEX, O.PRISM]
LC = {O.TRI: [[0, 0], [1, 0], [0, 1]], O.QUAD: [[0, 0], [1, 0], [1, 1], [0, 1]],
      O.TET: [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]],
      O.HEX: [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]],
      O.PRISM: [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1]]}


def _distorted(et, seed):
    rng = np.random.default_rng(seed)
    lc = np.asarray(LC[et], dtype=float)
    return lc * 1.2 + rng.uniform(-0.12, 0.12, size=lc.shape)


def _affine(et, seed):
    rng = np.random.default_rng(seed)
    dim = O.DIM[et]
    B = np.eye(dim) * 1.1 + rng.uniform(-0.2, 0.2, size=(dim, dim))
    x0 = rng.normal(size=dim)
    return np.asarray(LC[et], dtype=float) @ B.T + x0, B, x0


def _linear_state(co, A, b, c, p0):
    nsh, dim = co.shape
    u = np.zeros((nsh, dim + 1))
    u[:, :dim] = co @ A.T + b
    u[:, dim] = co @ c + p0
    return u


def _one(p, et, co, u, what, at, lips=None, uold=None, dt=0.0):
    conn = np.arange(O.NSH[et], dtype=np.int32)[None, :]
    return O.export(p, et, co, conn, u, what, at, lips, uold, dt)[0]


@pytest.mark.parametrize("et", ETS)
def test_linear_fields_at_scvf_ips(et):
    dim = O.DIM[et]
    rng = np.random.default_rng(1)
    co = _distorted(et, 3)
    A, b, c, p0 = rng.normal(size=(dim, dim)), rng.normal(size=dim), rng.normal(size=dim), 0.7
    u = _linear_state(co, A, b, c, p0)
    p = O.Params(dim=dim)
    G = O.geometry(et, co)
    ipg = G["ipg"]
    np.testing.assert_allclose(_one(p, et, co, u, "velocity", "scvf"), ipg @ A.T + b, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(_one(p, et, co, u, "pressure", "scvf")[:, 0], ipg @ c + p0, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(_one(p, et, co, u, "velocity_grad", "scvf"),
                               np.broadcast_to(A.reshape(-1), (O.NIP[et], dim * dim)), rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(_one(p, et, co, u, "pressure_grad", "scvf"),
                               np.broadcast_to(c, (O.NIP[et], dim)), rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("et", ETS)
def test_linear_fields_at_corners_and_local_points(et):
    dim = O.DIM[et]
    rng = np.random.default_rng(2)
    co, B, x0 = _affine(et, 4)
    A, b, c, p0 = rng.normal(size=(dim, dim)), rng.normal(size=dim), rng.normal(size=dim), -0.3
    u = _linear_state(co, A, b, c, p0)
    p = O.Params(dim=dim)
    # corners: nodal values
    np.testing.assert_allclose(_one(p, et, co, u, "velocity", "scv"), u[:, :dim], rtol=1e-13, atol=1e-14)
    assert np.array_equal(_one(p, et, co, u, "pressure", "scv")[:, 0], u[:, dim])
    # local points inside the reference element: x = B xi + x0
    xi = rng.uniform(0.05, 0.3, size=(5, dim))
    x = xi @ B.T + x0
    np.testing.assert_allclose(_one(p, et, co, u, "velocity", "local", xi), x @ A.T + b, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(_one(p, et, co, u, "velocity_ip", "local", xi), x @ A.T + b, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(_one(p, et, co, u, "pressure", "local", xi)[:, 0], x @ c + p0, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(_one(p, et, co, u, "pressure_grad", "local", xi),
                               np.broadcast_to(c, (5, dim)), rtol=1e-11, atol=1e-12)
    # the general case of ex_velocity_grad returns d u / d xi = A B
    np.testing.assert_allclose(_one(p, et, co, u, "velocity_grad", "local", xi),
                               np.broadcast_to((A @ B).reshape(-1), (5, dim * dim)), rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("et", ETS)
@pytest.mark.parametrize("cfg", [{"upwind": "full"}, {"upwind": "lps"}, {"upwind": "skewed", "difflen": "cor"},
                                 {"upwind": "full", "stokes": True}, {"upwind": "no", "timedep": True}])
def test_stabilised_ip_velocity_closes_continuity(et, cfg):
    """d(P, corner) of the element defect = sum over SCVFs of +-(v_stab.n) rho."""
    dim, nsh = O.DIM[et], O.NSH[et]
    cfg = dict(cfg)
    timedep = cfg.pop("timedep", False)
    p = O.Params(dim=dim, nu=0.02, rho=1.3, **cfg)
    co = _distorted(et, 7)
    rng = np.random.default_rng(8)
    u = rng.normal(size=(nsh, dim + 1))
    uold = rng.normal(size=(nsh, dim + 1)) if timedep else None
    dt = 0.05 if timedep else 0.0
    vs = _one(p, et, co, u, "velocity_ip", "scvf", None, uold, dt)
    G = O.geometry(et, co)
    want = np.zeros(nsh)
    for ip in range(O.NIP[et]):
        f = (vs[ip] @ G["n"][ip]) * p.rho
        want[G["from"][ip]] += f
        want[G["to"][ip]] -= f
    _, d = O.elem_A(p, et, co, u.reshape(-1), None if uold is None else uold.reshape(-1), dt)
    got = d.reshape(nsh, dim + 1)[:, dim]
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-13 * np.abs(want).max())


@pytest.mark.parametrize("et", ETS)
def test_unstabilised_ip_velocity_is_interpolated(et):
    dim, nsh = O.DIM[et], O.NSH[et]
    p = O.Params(dim=dim, stab="without")
    co = _distorted(et, 9)
    u = np.random.default_rng(3).normal(size=(nsh, dim + 1))
    np.testing.assert_allclose(_one(p, et, co, u, "velocity_ip", "scvf"), _one(p, et, co, u, "velocity", "scvf"),
                               rtol=1e-14, atol=1e-15)


def test_velocity_ip_needs_stabilisation():
    p = O.Params(dim=2, stab=None)
    co = _distorted(O.QUAD, 1)
    with pytest.raises(O.OracleError):
        _one(p, O.QUAD, co, np.zeros((4, 3)), "velocity_ip", "scvf")
