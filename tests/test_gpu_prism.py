"""Prisms through the C ABI against the oracle (generic path): the CSR
pattern bit-exactly, the A part in FP64 for every upwind, stabilisation and
flag case of the tetrahedral/hexahedral parity suite, the time-dependent
A + M + rhs assembly, the FP32 Jacobian bound, the no-allocation entry,
function-valued imports, and the refused combinations (the 3D regular
upwind, pyramids).

Tolerance: the parity rule (|a - b| <= 1e-12 |b| + 1e-13 S) plus an
absolute floor of 1e-15 max|b| over the array.  On prisms whose vertical
edges are axis-parallel (every element touching the unjittered boundary),
dx/dzeta at an ip is a sum of six products that cancels exactly in real
arithmetic; whether it comes out as 0 or as ~1e-17 depends on the order of
the sum, so an oracle entry that is exactly 0 (S = 0) can be 1e-19 here
(the kernel sums in its canonical corner order)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from plugin_navierstokes_amd import F32, F64, PART_A, PART_M, PART_RHS, NavierStokesFV1, NSFV1Error
from plugin_navierstokes_amd import meshes as M
from tests._parity import configure_disc, fp64_violation, oracle_params, worst
from tests._parity import run_case as _run
from tests.test_gpu_parity import CASES, _case_id

pytestmark = pytest.mark.gpu


FLOOR = 1e-15


def _viol(a, b, scale):
    a, b, scale = np.asarray(a, dtype=np.float64), np.asarray(b), np.asarray(scale)
    tol = 1e-12 * np.abs(b) + 1e-13 * scale + FLOOR * np.abs(b).max() + 1e-300
    return float(np.max(np.abs(a - b) / tol))


def _viol32(a, b, scale):
    a, b, scale = np.asarray(a, dtype=np.float64), np.asarray(b), np.asarray(scale)
    tol = 2.0 ** -17 * scale + 2.0 ** -20 * np.abs(b) + FLOOR * np.abs(b).max() + 1e-300
    return float(np.max(np.abs(a - b) / tol))


def _mesh(n=3, seed=3, jitter=0.1):
    return M.cavity_3d(n, M.PRISM, jitter, seed)


def test_pattern_bitexact(cuda):
    mesh = _mesh()
    disc = NavierStokesFV1("u,v,w,p", "Inner")
    disc.set_mesh(mesh)
    r, c = disc.pattern()
    r0, c0 = O.pattern(O.PRISM, mesh.nv, mesh.conn)
    assert np.array_equal(r, r0) and np.array_equal(c, c0)


@pytest.mark.parametrize("cfg", CASES, ids=_case_id)
def test_A_part_fp64(cuda, cfg):
    print(f"case prism cfg={cfg}", flush=True)
    J, d, Jr, dr, Ja, da, disc = _run(cuda, O.PRISM, cfg, mesh=_mesh())
    assert disc.assembly_path_used() == 1
    print(f"without the floor: J {fp64_violation(J, Jr, Ja):.3g}, d {fp64_violation(d, dr, da):.3g}; "
          f"max |J - J_oracle| {np.abs(J - Jr).max():.3g} of max |J| {np.abs(Jr).max():.3g}")
    assert _viol(J, Jr, Ja) <= 1.0, worst(J, Jr, Ja)
    assert _viol(d, dr, da) <= 1.0, worst(d, dr, da)


@pytest.mark.parametrize("up", ["no", "full", "skewed", "lps"])
def test_time_dependent_all_parts(cuda, up):
    cfg = {"upwind": up, "source": (0.1, 0.2, -0.3)}
    J, d, Jr, dr, Ja, da, _ = _run(cuda, O.PRISM, cfg, mesh=_mesh(seed=4), time_dep=True,
                                   parts=PART_A | PART_M | PART_RHS, s=(0.7, 1.3, -0.7))
    assert _viol(J, Jr, Ja) <= 1.0, worst(J, Jr, Ja)
    assert _viol(d, dr, da) <= 1.0, worst(d, dr, da)


@pytest.mark.parametrize("up", ["full", "lps"])
def test_fp32_jacobian_bound(cuda, up):
    J, d, Jr, dr, Ja, da, _ = _run(cuda, O.PRISM, {"upwind": up}, mesh=_mesh(seed=5), time_dep=True,
                                   parts=PART_A | PART_M, s=(1.0, 1.5, 0.0), jprec=32)
    assert _viol32(J, Jr, Ja) <= 1.0
    assert _viol(d, dr, da) <= 1.0, worst(d, dr, da)


def test_noalloc_and_imports(cuda):
    import torch

    mesh = _mesh(seed=6)
    disc = NavierStokesFV1("u,v,w,p", "Inner")
    configure_disc(disc, {"upwind": "lps"})
    disc.set_mesh(mesh)
    u = torch.from_numpy(M.random_state(mesh, 2)).to(cuda)
    J1, d1 = disc.assemble(u, jprec=F32)
    J2 = torch.empty_like(J1)
    d2 = torch.empty_like(d1)
    disc.assemble(u, jprec=F32, J=J2, d=d2, noalloc=True)
    assert torch.equal(J1, J2) and torch.equal(d1, d2)
    # function-valued viscosity on prisms, against the oracle
    nu = lambda x: 0.02 * (1.0 + x[:, 0])
    disc.set_kinematic_viscosity(nu)
    J, d = disc.assemble(u, jprec=F64)
    xo = np.stack([O.geometry(O.PRISM, mesh.xyz[mesh.conn[e]])["ipg"] for e in range(mesh.ne)])
    assert np.max(np.abs(disc.ip_positions("scvf") - xo)) <= 1e-14
    p = O.Params(dim=3, upwind="lps", stab="fields", nu=0.05, nu_ip=nu(xo.reshape(-1, 3)).reshape(mesh.ne, 9))
    rp, ci = disc.pattern()
    un = u.cpu().numpy()
    Jr, dr = O.assemble(p, O.PRISM, mesh.xyz, mesh.conn, un, rowptr=rp, colidx=ci)
    Ja, da = O.assemble_abs(p, O.PRISM, mesh.xyz, mesh.conn, un, rowptr=rp, colidx=ci)
    assert _viol(J.cpu().numpy(), Jr, Ja) <= 1.0
    assert _viol(d.cpu().numpy(), dr, da) <= 1.0


def test_refused_combinations(cuda):
    import torch

    mesh = _mesh()
    disc = NavierStokesFV1("u,v,w,p", "Inner")
    configure_disc(disc, {"upwind": "full"})
    disc.set_mesh(mesh)
    u = torch.from_numpy(M.random_state(mesh, 2)).to(cuda)
    # the coupled path runs on prisms (tests/test_gpu_coupled.py)
    disc.set_stabilization("flow")
    disc.assemble(u, jprec=F64)
    assert disc.assembly_path_used() == 4
    disc.set_stabilization("fields")
    # boundary discs run on prisms (tests/test_gpu_outflow.py)
    belem, bside = M.boundary_sides(mesh, mesh.boundary["lid"])
    assert len(belem) > 0
    disc.set_outflow(belem, bside)
    disc.assemble(u, jprec=F64)
    disc.set_outflow([], [])
    # exports run on prisms (tests/test_gpu_export.py), the coupled
    # configurations' stabilised ip velocity too
    assert disc.export(u, "velocity", "scv").shape == (mesh.ne, 6, 3)
    disc.set_upwind("positive")
    assert disc.export(u, "velocity_ip", "scvf").shape == (mesh.ne, 9, 3)
    # the regular upwind is 2D only (upwind.cpp:806)
    disc.set_upwind("regular")
    with pytest.raises(NSFV1Error) as e:
        disc.assemble(u, jprec=F64)
    assert e.value.name == "NOT_IMPLEMENTED"
    # pyramids are refused when the mesh is set
    pyr = NavierStokesFV1("u,v,w,p", "Inner")
    xyz = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0.5, 0.5, 1]], dtype=float)
    with pytest.raises(NSFV1Error) as e:
        pyr.set_mesh(et=5, xyz=xyz, conn=np.arange(5, dtype=np.int32)[None])
    assert e.value.name == "NOT_IMPLEMENTED"


def test_stabilised_velocity_export_covers_every_prism(cuda):
    """The stabilised ip velocity of every SCVF of a 128-prism mesh (one
    launch lane per (element, corner): six per prism) against the oracle."""
    import torch

    mesh = M.cavity_3d(4, M.PRISM, 0.1, 5)
    cfg = {"upwind": "full"}
    disc = NavierStokesFV1("u,v,w,p", "Inner")
    configure_disc(disc, cfg)
    disc.set_mesh(mesh)
    u = M.random_state(mesh, 7)
    got = disc.export(torch.from_numpy(u).to(cuda), "velocity_ip", "scvf").cpu().numpy()
    want = O.export(oracle_params(3, cfg), M.PRISM, mesh.xyz, mesh.conn, u, "velocity_ip", "scvf")
    assert got.shape == want.shape == (mesh.ne, 9, 3)
    tol = 1e-12 * np.abs(want) + 1e-12 * np.abs(want).max()
    assert np.max(np.abs(got - want) / tol) <= 1.0
