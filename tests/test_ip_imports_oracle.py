"""Pins the oracle's imports at the integration points
(navier_stokes_fv1.cpp:188-247): the kinematic viscosity and the source at
the SCVF ips, the source at the SCV ips.
- constant arrays reproduce the scalar path bitwise (every element type,
  FIELDS + LPS, time-dependent, mass and rhs parts);
- without stabilisation the element defect and Jacobian are affine in each
  ip's viscosity, D(nu) = D(0) + sum_ip nu_ip (D(e_ip) - D(0)), and the
  indicator of ip changes only the momentum rows of its two corners (the
  ip <-> edge mapping);
- with FIELDS the defect is affine in the SCVF source, the Jacobian does not
  depend on it, and (diagonal case) the indicator of (ip, d) changes only
  the continuity rows of the ip's two corners;
- the rhs part is f(corner) |SCV| rho per corner, with the SCV volumes of
  the geometry."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from plugin_navierstokes_amd import meshes as M

ETS = [O.TRI, O.QUAD, O.TET, O.HEX]
REFC = {
    O.TRI: [[0, 0], [1, 0.1], [0.2, 0.9]],
    O.QUAD: [[0, 0], [1, 0.05], [1.1, 1], [-0.05, 0.95]],
    O.TET: [[0, 0, 0], [1, 0.1, 0], [0.1, 0.9, 0.1], [0.05, 0.1, 1]],
    O.HEX: [[0, 0, 0], [1, 0, 0.05], [1, 1, 0], [0, 1, 0], [0, 0.05, 1], [1, 0, 1], [1.05, 1, 1], [0, 1, 1.1]],
}


def _mesh(et):
    if O.DIM[et] == 2:
        return M.cavity_2d(3, M.TRI if et == O.TRI else M.QUAD, 0.1, seed=1)
    return M.cavity_3d(2, M.TET if et == O.TET else M.HEX, 0.1, seed=1)


@pytest.mark.parametrize("et", ETS)
def test_constant_arrays_equal_the_scalar_path(et):
    mesh = _mesh(et)
    dim, ne, nip, nsh = O.DIM[et], mesh.ne, O.NIP[et], O.NSH[et]
    u = M.random_state(mesh, 3)
    uold = M.random_state(mesh, 4)
    f = (0.3, -0.2, 0.1)[:dim]
    p = O.Params(dim=dim, upwind="lps", stab="fields", nu=0.02, source=f)
    q = O.Params(dim=dim, upwind="lps", stab="fields", nu=0.02, source=f, nu_ip=np.full((ne, nip), 0.02),
                 f_scv=np.tile(f, (ne, nsh, 1)), f_scvf=np.tile(f, (ne, nip, 1)))
    J1, d1 = O.assemble(p, et, mesh.xyz, mesh.conn, u, uold, 0.01, s_m=1.0, s_f=-1.0, parts=7)
    J2, d2 = O.assemble(q, et, mesh.xyz, mesh.conn, u, uold, 0.01, s_m=1.0, s_f=-1.0, parts=7)
    assert np.array_equal(J1, J2) and np.array_equal(d1, d2)


def _elem(et, seed):
    rng = np.random.default_rng(seed)
    co = np.array(REFC[et], dtype=float)
    ul = rng.normal(size=(O.NSH[et], O.DIM[et] + 1))
    return co, ul


@pytest.mark.parametrize("et", ETS)
def test_viscosity_per_ip_is_affine_and_local(et):
    co, ul = _elem(et, 5)
    dim, nip, nsh, nc = O.DIM[et], O.NIP[et], O.NSH[et], O.DIM[et] + 1
    G = O.geometry(et, co)

    def A(nu):
        return O.elem_A(O.Params(dim=dim, upwind="full", stab="without", nu=1.0, nu_ip=nu[None, :]), et, co, ul)

    J0, d0 = A(np.zeros(nip))
    dJ, dd = [], []
    for ip in range(nip):
        e = np.zeros(nip)
        e[ip] = 1.0
        J, d = A(e)
        dJ.append(J - J0)
        dd.append(d - d0)
        rows = np.zeros((nsh, nc), dtype=bool)
        rows[[G["from"][ip], G["to"][ip]], :dim] = True
        rows = rows.ravel()
        assert np.all(dd[-1][~rows] == 0) and np.all(dJ[-1][~rows] == 0), ip
        assert np.any(dd[-1][rows] != 0) and np.any(dJ[-1][rows] != 0), ip
        # the from and to rows receive opposite fluxes (conservation)
        fr, to = G["from"][ip], G["to"][ip]
        np.testing.assert_allclose(dd[-1].reshape(nsh, nc)[fr], -dd[-1].reshape(nsh, nc)[to], rtol=1e-12, atol=1e-14)
    nu = np.random.default_rng(6).uniform(0.01, 0.1, nip)
    J, d = A(nu)
    Jl = J0 + sum(nu[i] * dJ[i] for i in range(nip))
    dl = d0 + sum(nu[i] * dd[i] for i in range(nip))
    np.testing.assert_allclose(d, dl, rtol=1e-12, atol=1e-13 * np.abs(d).max())
    np.testing.assert_allclose(J, Jl, rtol=1e-12, atol=1e-13 * np.abs(J).max())


@pytest.mark.parametrize("et", ETS)
@pytest.mark.parametrize("upw", ["full", "lps"])
def test_scvf_source_is_affine_and_local(et, upw):
    co, ul = _elem(et, 7)
    dim, nip, nsh, nc = O.DIM[et], O.NIP[et], O.NSH[et], O.DIM[et] + 1
    G = O.geometry(et, co)
    uold = ul + 0.01

    def A(fs):
        p = O.Params(dim=dim, upwind=upw, stab="fields", nu=0.05, source=(0.0,) * dim, f_scvf=fs[None])
        return O.elem_A(p, et, co, ul, uold, 0.02)

    J0, d0 = A(np.zeros((nip, dim)))
    dd = {}
    for ip in range(nip):
        for k in range(dim):
            e = np.zeros((nip, dim))
            e[ip, k] = 1.0
            J, d = A(e)
            assert np.array_equal(J, J0)  # the source is a constant of the linearisation
            dd[ip, k] = d - d0
            if upw == "full":  # diagonal case: the ip's own stabilised velocity only
                rows = np.zeros((nsh, nc), dtype=bool)
                rows[[G["from"][ip], G["to"][ip]], dim] = True
                assert np.all(dd[ip, k][~rows.ravel()] == 0), (ip, k)
                assert np.any(dd[ip, k][rows.ravel()] != 0), (ip, k)
    fs = np.random.default_rng(8).normal(size=(nip, dim))
    _, d = A(fs)
    dl = d0 + sum(fs[ip, k] * dd[ip, k] for ip in range(nip) for k in range(dim))
    np.testing.assert_allclose(d, dl, rtol=1e-10, atol=1e-12 * np.abs(d).max())


@pytest.mark.parametrize("et", ETS)
def test_scv_source_closed_form(et):
    co, _ = _elem(et, 9)
    dim, nsh, nc = O.DIM[et], O.NSH[et], O.DIM[et] + 1
    fv = np.random.default_rng(10).normal(size=(nsh, dim))
    rho = 1.3
    f = O.elem_rhs(O.Params(dim=dim, rho=rho, source=(0.0,) * dim, f_scv=fv[None]), et, co).reshape(nsh, nc)
    vol = O.geometry(et, co)["vol"]
    np.testing.assert_allclose(f[:, :dim], fv * vol[:, None] * rho, rtol=1e-15)
    assert np.all(f[:, dim] == 0)
