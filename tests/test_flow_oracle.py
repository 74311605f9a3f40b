"""Pins the oracle's positive upwind (upwind.cpp:643-786), the matrix case
of the FIELDS stabilisation (stabilization.cpp:244-403) and the FLOW
stabilisation (stabilization.cpp:436-772):

- positive upwind shapes against a hand-derived case (diagonal flow through
  the unit square) and its invariants (consistency sum U + sum W = 1,
  positivity, the half/half rule without flux);
- the matrix solve against the diagonal formulas where both apply (upwind
  without ip shapes, forced through the matrix path);
- FLOW = FIELDS where the velocity vanishes (no convection, no cross terms);
- reproduction of a constant velocity by FIELDS and FLOW with the positive
  upwind (only true if the ip shapes enter with the right sign and index);
- the Picard identity d = J u and conservation of the assembled defect, and
  the continuity closure of the stabilised ip velocity;
- the matrix case against numpy.linalg.solve of the ip system built from the
  upwind's shapes on unit elements (closed-form diffusion length)."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from plugin_navierstokes_amd import meshes as M

ETS = [O.TRI, O.QUAD, O.TET, O.HEX, O.PRISM]
LC = {O.TRI: [[0, 0], [1, 0], [0, 1]], O.QUAD: [[0, 0], [1, 0], [1, 1], [0, 1]],
      O.TET: [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]],
      O.HEX: [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]],
      O.PRISM: [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1]]}


def _elem(et, seed, jitter=0.12):
    rng = np.random.default_rng(seed)
    lc = np.asarray(LC[et], dtype=float)
    return lc * 1.1 + rng.uniform(-jitter, jitter, size=lc.shape)


def _ipvel(et, co, u):
    g = O.geometry(et, co)
    return g, g["phi"] @ u[:, : O.DIM[et]]


def test_positive_upwind_diagonal_flow_unit_square():
    """v = (1, 1) on the unit square: SCV 0 has only outflow (ip0, ip3 take
    corner 0); SCV 1 passes the inflow through ip0 on to ip1 (W[1,0] = 1);
    SCV 3 passes ip3 on to ip2 (W[2,3] = 1)."""
    co = np.asarray(LC[O.QUAD], dtype=float)
    v = np.ones((4, 2))
    U, W, L = O.upwind("positive", O.QUAD, co, v, ip_shapes=True)
    g = O.geometry(O.QUAD, co)
    assert [tuple(e) for e in zip(g["from"], g["to"])] == [(0, 1), (1, 2), (2, 3), (3, 0)]
    wantU = np.zeros((4, 4))
    wantU[0, 0] = 1.0
    wantU[3, 0] = 1.0
    wantW = np.zeros((4, 4))
    wantW[1, 0] = 1.0
    wantW[2, 3] = 1.0
    np.testing.assert_allclose(U, wantU, atol=1e-15)
    np.testing.assert_allclose(W, wantW, atol=1e-15)
    ipg = g["ipg"]
    np.testing.assert_allclose(L[1], np.linalg.norm(ipg[1] - ipg[0]), rtol=1e-14)
    np.testing.assert_allclose(L[0], np.linalg.norm(ipg[0] - co[0]), rtol=1e-14)


@pytest.mark.parametrize("et", ETS)
@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_positive_upwind_invariants(et, seed):
    co = _elem(et, seed)
    rng = np.random.default_rng(seed + 10)
    u = rng.normal(size=(O.NSH[et], O.DIM[et] + 1))
    g, v = _ipvel(et, co, u)
    U, W, L = O.upwind("positive", et, co, v, ip_shapes=True)
    np.testing.assert_allclose(U.sum(1) + W.sum(1), 1.0, rtol=1e-14)
    assert U.min() >= -1e-15 and W.min() >= 0.0
    assert np.all(np.diag(W) == 0.0)
    # an ip only takes values from the SCV it leaves: corner `from` if the
    # flux is positive, else `to`
    for ip in range(O.NIP[et]):
        fl = v[ip] @ g["n"][ip]
        up = g["from"][ip] if fl > 0 else g["to"][ip]
        others = [k for k in range(O.NSH[et]) if k != up]
        assert np.all(U[ip, others] == 0.0)
    # the upwind point is a convex combination: conv length > 0
    assert np.all(L > 0)


def test_positive_upwind_without_flux():
    et = O.HEX
    co = _elem(et, 4)
    U, W, _ = O.upwind("positive", et, co, np.zeros((12, 3)), ip_shapes=True)
    g = O.geometry(et, co)
    for ip in range(12):
        assert U[ip, g["from"][ip]] == 0.5 and U[ip, g["to"][ip]] == 0.5
    assert not W.any()


def _rand_u(et, seed):
    return np.random.default_rng(seed).normal(size=(O.NSH[et], O.DIM[et] + 1))


@pytest.mark.parametrize("et", ETS)
@pytest.mark.parametrize("stab", ["fields", "flow"])
@pytest.mark.parametrize("upwind", ["full", "lps"])
@pytest.mark.parametrize("timedep", [False, True])
def test_matrix_solve_equals_diagonal_case(et, stab, upwind, timedep):
    """Without ip shapes the ip system is diagonal; forcing it through the
    matrix path must give the same stabilised velocity, shapes and element
    matrices."""
    co = _elem(et, 5)
    u = _rand_u(et, 6)
    uold = _rand_u(et, 7) if timedep else None
    dt = 0.03 if timedep else 0.0
    src = (0.2, -0.1, 0.3)[: O.DIM[et]]
    base = dict(dim=O.DIM[et], upwind=upwind, stab=stab, nu=0.04, rho=1.2, source=src, pac=False)
    for pac in (False, True):
        base["pac"] = pac
        Jd, dd = O.elem_A(O.Params(**base), et, co, u.ravel(), None if uold is None else uold.ravel(), dt)
        Jm, dm = O.elem_A(O.Params(force_matrix=True, **base), et, co, u.ravel(),
                          None if uold is None else uold.ravel(), dt)
        np.testing.assert_allclose(Jm, Jd, rtol=1e-11, atol=1e-12 * np.abs(Jd).max())
        np.testing.assert_allclose(dm, dd, rtol=1e-11, atol=1e-12 * np.abs(dd).max())


@pytest.mark.parametrize("et", ETS)
def test_flow_equals_fields_without_velocity(et):
    co = _elem(et, 8)
    u = _rand_u(et, 9)
    u[:, : O.DIM[et]] = 0.0
    for up in ("full", "positive"):
        pf = O.Params(dim=O.DIM[et], upwind=up, stab="fields", nu=0.02)
        pw = O.Params(dim=O.DIM[et], upwind=up, stab="flow", nu=0.02)
        Jf, df = O.elem_A(pf, et, co, u.ravel())
        Jw, dw = O.elem_A(pw, et, co, u.ravel())
        np.testing.assert_allclose(dw, df, rtol=1e-13, atol=1e-14)
        # the Jacobians differ only by FLOW's cross-component shapes, which
        # carry the (zero) ip velocity
        np.testing.assert_allclose(Jw, Jf, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("et", ETS)
@pytest.mark.parametrize("stab", ["fields", "flow"])
@pytest.mark.parametrize("timedep", [False, True])
def test_constant_velocity_reproduced_with_positive_upwind(et, stab, timedep):
    """A constant velocity with constant pressure is reproduced by the
    stabilised ip velocity: the continuity rows are sum +-(U.n)."""
    co = _elem(et, 11)
    dim = O.DIM[et]
    Uc = np.random.default_rng(12).uniform(-1, 1, dim)
    u = np.zeros((O.NSH[et], dim + 1))
    u[:, :dim] = Uc
    u[:, dim] = 0.4
    p = O.Params(dim=dim, upwind="positive", stab=stab, nu=0.03)
    uold, dt = (u.copy(), 0.02) if timedep else (None, 0.0)
    vs = O.export(p, et, co, np.arange(O.NSH[et], dtype=np.int32)[None], u, "velocity_ip", "scvf", None, uold, dt)[0]
    np.testing.assert_allclose(vs, np.broadcast_to(Uc, vs.shape), rtol=1e-12, atol=1e-13)


CFGS = [dict(upwind="positive"), dict(upwind="positive", stab="flow"), dict(upwind="full", stab="flow"),
        dict(upwind="lps", stab="flow", difflen="cor"), dict(upwind="skewed", stab="flow"),
        dict(upwind="positive", pac=True), dict(upwind="positive", stab="flow", pac=True),
        dict(upwind="full", stab_upwind="positive"), dict(upwind="positive", stab_upwind="full"),
        dict(upwind="no", stab="flow", stokes=True)]


@pytest.mark.parametrize("et", ETS)
@pytest.mark.parametrize("cfg", CFGS, ids=lambda c: "-".join(f"{k}={v}" for k, v in c.items()))
def test_picard_identity_and_conservation(et, cfg):
    mesh = (M.cavity_2d(4, et, 0.1, 3) if et in (O.TRI, O.QUAD) else M.cavity_3d(2, et, 0.1, 3))
    u = M.random_state(mesh, 5)
    p = O.Params(dim=mesh.dim, nu=0.05, **cfg)
    rp, ci = O.pattern(et, mesh.nv, mesh.conn)
    J, d = O.assemble(p, et, mesh.xyz, mesh.conn, u, rowptr=rp, colidx=ci)
    A = sp.csr_matrix((J, ci, rp), shape=(len(rp) - 1,) * 2)
    assert np.abs(A @ u.ravel() - d).max() <= 1e-12 * np.abs(d).max()
    assert np.abs(d.reshape(-1, mesh.dim + 1).sum(0)).max() <= 1e-12 * np.abs(d).max()


@pytest.mark.parametrize("et", ETS)
@pytest.mark.parametrize("cfg", [dict(upwind="positive"), dict(upwind="positive", stab="flow"),
                                 dict(upwind="skewed", stab="flow")])
def test_stabilised_velocity_closes_continuity(et, cfg):
    co = _elem(et, 13)
    dim, nsh = O.DIM[et], O.NSH[et]
    p = O.Params(dim=dim, nu=0.02, rho=1.3, **cfg)
    u = _rand_u(et, 14)
    vs = O.export(p, et, co, np.arange(nsh, dtype=np.int32)[None], u, "velocity_ip", "scvf")[0]
    g = O.geometry(et, co)
    want = np.zeros(nsh)
    for ip in range(O.NIP[et]):
        f = (vs[ip] @ g["n"][ip]) * p.rho
        want[g["from"][ip]] += f
        want[g["to"][ip]] -= f
    _, d = O.elem_A(p, et, co, u.ravel())
    np.testing.assert_allclose(d.reshape(nsh, dim + 1)[:, dim], want, rtol=1e-12, atol=1e-13 * np.abs(want).max())


@pytest.mark.parametrize("et,L2inv", [(O.QUAD, 32 / 7), (O.HEX, 3.2)])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_matrix_case_against_numpy_solve(et, L2inv, seed):
    """FIELDS with the positive upwind on the unit square / cube (RAW
    diffusion length in closed form): assemble the ip system from the
    upwind's corner and ip shapes and solve it with numpy.linalg.solve;
    the oracle's stabilised velocities must agree (stabilization.cpp:
    244-403, steady, no source)."""
    co = np.asarray(LC[et], dtype=float)
    dim, nsh, nip = O.DIM[et], O.NSH[et], O.NIP[et]
    rng = np.random.default_rng(30 + seed)
    u = rng.normal(size=(nsh, dim + 1))
    nu, rho = 0.04, 1.3
    g = O.geometry(et, co)
    sv = g["phi"] @ u[:, :dim]
    U, W, L = O.upwind("positive", et, co, sv, ip_shapes=True)
    vv = nu * L2inv
    vc = np.linalg.norm(sv, axis=1) / L
    Mx = np.diag(vv + vc) - vc[:, None] * W
    rhs = np.zeros((nip, dim))
    for ip in range(nip):
        cont = vv * g["phi"][ip] + vc[ip] * U[ip]
        for d in range(dim):
            rhs[ip, d] = cont @ u[:, d] - (g["grad"][ip][:, d] / rho) @ u[:, dim]
    want = np.linalg.solve(Mx, rhs)
    p = O.Params(dim=dim, upwind="positive", nu=nu, rho=rho)
    got = O.export(p, et, co, np.arange(nsh, dtype=np.int32)[None], u, "velocity_ip", "scvf")[0]
    np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-12 * np.abs(want).max())
