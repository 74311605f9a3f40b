"""GPU parity: the HIP assembly (through the C ABI) against the CPU oracle on
the same seeded meshes and states, every element type, upwind,
stabilisation and flag of the reference path
(incompressible/fv1/navier_stokes_fv1.cpp:250-869, upwind.cpp:52-575,
incompressible/fv1/stabilization.cpp:102-242,805-850,
incompressible/fv1/diffusion_length.h:48-199)."""
from __future__ import annotations

import itertools

import numpy as np
import pytest

import oracle as O
from plugin_navierstokes_amd import NavierStokesFV1, NSFV1Error, PART_A, PART_M, PART_RHS, F32, F64
from plugin_navierstokes_amd import meshes as M
from tests._parity import (configure_disc, fp32_violation, fp64_violation, max_rel_err, oracle_params,
                           small_mesh, worst, FP32_REL, FP32_ABS)
from tests._parity import run_case as _run

pytestmark = pytest.mark.gpu

ETS = [M.TRI, M.QUAD, M.TET, M.HEX]
ET_NAMES = {M.TRI: "tri", M.QUAD: "quad", M.TET: "tet", M.HEX: "hex"}


@pytest.mark.parametrize("et", ETS, ids=lambda e: ET_NAMES[e])
def test_pattern_bitexact(cuda, et):
    mesh = small_mesh(et)
    disc = NavierStokesFV1("u,v,p" if mesh.dim == 2 else "u,v,w,p", "Inner")
    disc.set_mesh(mesh)
    r, c = disc.pattern()
    r0, c0 = O.pattern(et, mesh.nv, mesh.conn)
    assert np.array_equal(r, r0) and np.array_equal(c, c0)


CASES = []
for up in ["no", "full", "skewed", "lps"]:
    CASES.append({"upwind": up})
CASES += [
    {"upwind": "full", "difflen": "fivepoint"},
    {"upwind": "skewed", "difflen": "cor"},
    {"upwind": "lps", "stab": "without"},
    {"upwind": "full", "stokes": True},
    {"upwind": "skewed", "laplace": True},
    {"upwind": "full", "peclet": True},
    {"upwind": "lps", "exact": 1.0},
    {"upwind": "full", "exact": 0.5, "peclet": True},
    {"upwind": "skewed", "pac": True},
    {"upwind": "full", "pac": True, "exact": 1.0},
    {"upwind": "full", "stab_upwind": "skewed"},
    {"upwind": "lps", "rho": 1.7, "nu": 0.003, "source": (0.3, -0.2, 0.1)},
]


def _case_id(c):
    return "-".join(f"{k}={v}" for k, v in c.items())


@pytest.mark.parametrize("cfg", CASES, ids=_case_id)
@pytest.mark.parametrize("et", ETS, ids=lambda e: ET_NAMES[e])
def test_A_part_fp64(cuda, et, cfg):
    print(f"case et={ET_NAMES[et]} cfg={cfg}", flush=True)
    J, d, Jr, dr, Ja, da, _ = _run(cuda, et, cfg)
    assert fp64_violation(J, Jr, Ja) <= 1.0, worst(J, Jr, Ja)
    assert fp64_violation(d, dr, da) <= 1.0, worst(d, dr, da)


@pytest.mark.parametrize("et", ETS, ids=lambda e: ET_NAMES[e])
@pytest.mark.parametrize("up", ["full", "skewed", "lps"])
def test_time_dependent_fp64(cuda, et, up):
    cfg = {"upwind": up, "source": (0.1, 0.2, -0.3)}
    J, d, Jr, dr, Ja, da, _ = _run(cuda, et, cfg, time_dep=True, parts=PART_A | PART_M | PART_RHS,
                                   s=(0.7, 1.3, -0.7))
    assert fp64_violation(J, Jr, Ja) <= 1.0, worst(J, Jr, Ja)
    assert fp64_violation(d, dr, da) <= 1.0, worst(d, dr, da)


@pytest.mark.parametrize("et", ETS, ids=lambda e: ET_NAMES[e])
def test_mass_and_rhs_parts(cuda, et):
    cfg = {"upwind": "full", "source": (1.0, -2.0, 0.5), "rho": 1.3}
    for parts, s in [(PART_M, (0.0, 1.0, 0.0)), (PART_RHS, (0.0, 0.0, 1.0))]:
        J, d, Jr, dr, Ja, da, _ = _run(cuda, et, cfg, parts=parts, s=s)
        assert fp64_violation(J, Jr, Ja) <= 1.0, worst(J, Jr, Ja)
        assert fp64_violation(d, dr, da) <= 1.0, worst(d, dr, da)


@pytest.mark.parametrize("et", ETS, ids=lambda e: ET_NAMES[e])
@pytest.mark.parametrize("up", ["full", "skewed", "lps"])
def test_fp32_jacobian_bound(cuda, et, up):
    J, d, Jr, dr, Ja, da, _ = _run(cuda, et, {"upwind": up}, jprec=F32)
    assert J.dtype == np.float32
    assert fp32_violation(J, Jr, Ja) <= 1.0, worst(J, Jr, Ja, FP32_REL, FP32_ABS)
    assert fp64_violation(d, dr, da) <= 1.0, worst(d, dr, da)


@pytest.mark.parametrize("et", ETS, ids=lambda e: ET_NAMES[e])
def test_bitwise_repeatable(cuda, et):
    import torch

    mesh = small_mesh(et, n=6)
    disc = NavierStokesFV1("u,v,p" if mesh.dim == 2 else "u,v,w,p", "Inner")
    configure_disc(disc, {"upwind": "skewed"})
    disc.set_mesh(mesh)
    u = torch.from_numpy(M.random_state(mesh, 9)).to(cuda)
    J1, d1 = disc.assemble(u, jprec=F32)
    J2, d2 = disc.assemble(u, jprec=F32)
    assert torch.equal(J1, J2) and torch.equal(d1, d2)


def test_cavity_configs_medium(cuda):
    """Jittered lid-driven cavity states of the 2D configurations."""
    import torch

    for et, up, n in [(M.QUAD, "full", 16), (M.TRI, "lps", 16)]:
        mesh = M.cavity_2d(n, et, 0.1, seed=11)
        u = M.cavity_lid_state(mesh, 1.0, base=0.2 * M.random_state(mesh, 3))
        cfg = {"upwind": up, "nu": 1e-2 if et == M.QUAD else 1e-3}
        J, d, Jr, dr, Ja, da, _ = _run(cuda, et, cfg, mesh=mesh)
        assert fp64_violation(J, Jr, Ja) <= 1.0, worst(J, Jr, Ja)
        assert fp64_violation(d, dr, da) <= 1.0, worst(d, dr, da)


def test_error_cases(cuda):
    import torch

    mesh = small_mesh(M.QUAD)
    disc = NavierStokesFV1("u,v,p", "Inner")
    disc.set_mesh(mesh)
    u = torch.zeros(mesh.nv, 3, dtype=torch.float64, device=cuda)
    with pytest.raises(NSFV1Error) as e:
        disc.assemble(u)
    assert e.value.name == "NO_STABILIZATION"
    disc.set_stabilization("fields")
    with pytest.raises(NSFV1Error) as e:
        disc.assemble(u)
    assert e.value.name == "NO_UPWIND"
    disc.set_upwind("full")
    with pytest.raises(NSFV1Error) as e:
        disc.assemble(u)
    assert e.value.name == "NO_VISCOSITY"
    disc.set_kinematic_viscosity(0.1)
    disc.assemble(u)


@pytest.mark.parametrize("et", [M.HEX, M.TET, M.QUAD, M.TRI])
def test_upwind_record_replay(cuda, et):
    """Replaying the decisions recorded at u reproduces the fresh assembly at
    u bitwise; replay at another state uses the recorded corners."""
    import torch

    from plugin_navierstokes_amd import UPWIND_FRESH, UPWIND_RECORD, UPWIND_REPLAY

    mesh = M.cavity_2d(6, et, 0.1, 4) if et in (M.QUAD, M.TRI) else M.cavity_3d(4, et, 0.1, 4)
    disc = NavierStokesFV1("u,v,p" if mesh.dim == 2 else "u,v,w,p", "Inner")
    disc.set_stabilization("fields")
    disc.set_upwind("skewed")
    disc.set_kinematic_viscosity(0.01)
    disc.set_mesh(mesh)
    u = torch.from_numpy(M.random_state(mesh, 7)).to(cuda)
    uold = 0.9 * u
    J0, d0 = disc.assemble(u, uold, 0.01, jprec=F64)
    disc.set_upwind_decisions(UPWIND_RECORD)
    J1, d1 = disc.assemble(u, uold, 0.01, jprec=F64)
    disc.set_upwind_decisions(UPWIND_REPLAY)
    J2, d2 = disc.assemble(u, uold, 0.01, jprec=F64)
    assert torch.equal(J0, J1) and torch.equal(d0, d1)
    assert torch.equal(J0, J2) and torch.equal(d0, d2)
    # a reversed flow field replayed with the old decisions differs from a fresh assembly
    J3, d3 = disc.assemble(-u, -uold, 0.01, jprec=F64)
    disc.set_upwind_decisions(UPWIND_FRESH)
    J4, d4 = disc.assemble(-u, -uold, 0.01, jprec=F64)
    assert not torch.equal(d3, d4)
    disc.sync_errors()


def _affine_mesh(et, shear):
    """Unjittered structured mesh, optionally mapped by a fixed linear map:
    every quadrilateral / hexahedron is a parallelogram / parallelepiped, so
    the kernel's affine ray cast runs (simplices always take it)."""
    mesh = M.cavity_2d(5, et) if et in (M.TRI, M.QUAD) else M.cavity_3d(3, et)
    if shear:
        A = np.array([[1.0, 0.3], [-0.1, 1.2]]) if mesh.dim == 2 else \
            np.array([[1.0, 0.3, 0.1], [0.0, 1.1, -0.2], [0.05, 0.0, 0.9]])
        mesh.xyz = mesh.xyz @ A.T
    return mesh


@pytest.mark.parametrize("shear", [False, True], ids=["unit", "sheared"])
@pytest.mark.parametrize("up", ["skewed", "lps"])
@pytest.mark.parametrize("et", ETS, ids=lambda e: ET_NAMES[e])
def test_affine_elements_ray_cast(cuda, et, up, shear):
    print(f"case et={ET_NAMES[et]} up={up} shear={shear}", flush=True)
    mesh = _affine_mesh(et, shear)
    for jprec in (F64, F32):
        J, d, Jr, dr, Ja, da, _ = _run(cuda, et, {"upwind": up}, mesh=mesh, seed=11, time_dep=True,
                                       parts=PART_A | PART_M, s=(1.0, 2.0, 0.0), jprec=jprec)
        if jprec == F64:
            assert fp64_violation(J, Jr, Ja) <= 1.0, worst(J, Jr, Ja)
        else:
            assert fp32_violation(J, Jr, Ja) <= 1.0, worst(J, Jr, Ja, FP32_REL, FP32_ABS)
        assert fp64_violation(d, dr, da) <= 1.0, worst(d, dr, da)


@pytest.mark.parametrize("et", [M.TRI, M.QUAD, M.HEX], ids=lambda e: ET_NAMES[e])
@pytest.mark.parametrize("up", ["no", "full", "skewed", "lps"])
def test_split_row_lanes(cuda, et, up):
    """The row-lane gather (path 2) against the oracle and against the split
    gather (path 0), with the mass and source parts: the same terms in the
    same order per entry; the compiler may contract different products into
    FMAs, so the paths agree to FP32 rounding, not bitwise."""
    cfg = {"upwind": up, "source": (0.2, -0.1, 0.3)}
    mesh = small_mesh(et, n=6, seed=8)
    res = {}
    for path in (0, 2):
        J, d, Jr, dr, Ja, da, disc = _run(cuda, et, cfg, mesh=mesh, seed=4, time_dep=True,
                                              parts=PART_A | PART_M | PART_RHS, s=(0.8, 1.5, -0.5), jprec=F32,
                                              path=path)
        assert fp32_violation(J, Jr, Ja) <= 1.0, worst(J, Jr, Ja, FP32_REL, FP32_ABS)
        assert fp64_violation(d, dr, da) <= 1.0, worst(d, dr, da)
        res[path] = (J, d, Ja, da)
    J0, d0, Ja, da = res[0]
    J2, d2, _, _ = res[2]
    assert np.max(np.abs(J2.astype(np.float64) - J0) / (2.0 ** -20 * (Ja + np.abs(J0)) + 1e-300)) <= 1.0
    assert fp64_violation(d2, d0, da) <= 1.0


@pytest.mark.parametrize("path", [0, 2])
@pytest.mark.parametrize("et", [M.QUAD, M.HEX], ids=lambda e: ET_NAMES[e])
def test_split_defect_only(cuda, et, path):
    """A defect-only FP32 assembly takes the split path too: its defect is
    bitwise the one assembled together with the Jacobian, and matches the
    oracle."""
    import torch

    mesh = small_mesh(et, n=6, seed=8)
    cfg = {"upwind": "skewed"}
    disc = NavierStokesFV1("u,v,p" if mesh.dim == 2 else "u,v,w,p", "Inner")
    configure_disc(disc, cfg)
    disc.set_mesh(mesh)
    disc.set_assembly_path(path)
    uh, uoh = M.random_state(mesh, 5), M.random_state(mesh, 6)
    u, uo = torch.from_numpy(uh).to(cuda), torch.from_numpy(uoh).to(cuda)
    _, d_with = disc.assemble(u, uo, 0.02, s_m=1.0, parts=PART_A | PART_M, jprec=F32)
    _, d_only = disc.assemble(u, uo, 0.02, s_m=1.0, parts=PART_A | PART_M, jprec=F32, want_J=False)
    assert torch.equal(d_with, d_only)
    p = oracle_params(mesh.dim, cfg)
    _, dr = O.assemble(p, et, mesh.xyz, mesh.conn, uh, uoh, 0.02, s_m=1.0, parts=PART_A | PART_M, want_J=False)
    _, da = O.assemble_abs(p, et, mesh.xyz, mesh.conn, uh, uoh, 0.02, s_m=1.0, parts=PART_A | PART_M)
    assert fp64_violation(d_only.cpu().numpy(), dr, da) <= 1.0, worst(d_only.cpu().numpy(), dr, da)
