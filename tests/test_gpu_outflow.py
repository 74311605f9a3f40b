"""GPU parity of the FV1 boundary discs -- no-normal-stress outflow
(incompressible/fv1/bnd/no_normal_stress_outflow_fv1.cpp:188-426),
symmetry (incompressible/fv1/bnd/symmetric_boundary_fv1.cpp:224-386) and
sliding wall (incompressible/fv1/bnd/wall_sliding_fv1.cpp:229-425):
element assembly plus boundary rows through the C ABI against the CPU
oracle."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from plugin_navierstokes_amd import F32, F64, NavierStokesFV1, PART_A
from plugin_navierstokes_amd import meshes as M
from tests._parity import configure_disc, oracle_params

pytestmark = pytest.mark.gpu

ETS = [M.TRI, M.QUAD, M.TET, M.HEX, M.PRISM]
# prisms: the absolute floor of their parity suite (an entry that cancels
# exactly in the oracle's corner order can be ~1e-19 in the kernel's order)
PRISM_FLOOR = 1e-15


def _case(et):
    mesh = M.cavity_2d(5, et, 0.12, 8) if et in (M.TRI, M.QUAD) else M.cavity_3d(3, et, 0.12, 8)
    x = mesh.xyz[:, 0]
    out_nodes = np.nonzero(np.isclose(x, x.max()))[0]
    top = np.nonzero(np.isclose(mesh.xyz[:, 1], mesh.xyz[:, 1].max()))[0]
    be1, bs1 = M.boundary_sides(mesh, out_nodes)
    be2, bs2 = M.boundary_sides(mesh, top)
    return mesh, np.concatenate([be1, be2]), np.concatenate([bs1, bs2])


def _left(mesh):
    """Sides on x = min (the sliding wall of the combined case)."""
    x = mesh.xyz[:, 0]
    return M.boundary_sides(mesh, np.nonzero(np.isclose(x, x.min()))[0])


def _bottom(mesh):
    """Sides on y = min (the symmetry plane of the combined case)."""
    y = mesh.xyz[:, 1]
    return M.boundary_sides(mesh, np.nonzero(np.isclose(y, y.min()))[0])


@pytest.mark.parametrize("cfg", [{"upwind": "full"}, {"upwind": "lps", "laplace": True}, {"upwind": "full", "stokes": True}],
                         ids=["full", "lps-laplace", "stokes"])
@pytest.mark.parametrize("jprec", [F64, F32])
@pytest.mark.parametrize("et", ETS)
@pytest.mark.parametrize("kind", ["outflow", "symmetry", "both", "sliding", "all"])
def test_outflow_parity(cuda, et, jprec, cfg, kind):
    import torch

    mesh, belem, bside = _case(et)
    assert len(belem) > 0
    u = M.random_state(mesh, 21)
    disc = NavierStokesFV1("u,v,p" if mesh.dim == 2 else "u,v,w,p", "Inner")
    configure_disc(disc, cfg)
    disc.set_mesh(mesh)
    rp, ci = disc.pattern()
    p = oracle_params(mesh.dim, cfg)
    Jo, do = 0.0, 0.0
    if kind in ("outflow", "both", "all"):
        disc.set_outflow(belem, bside)
        Jx, dx = O.outflow(p, et, mesh.xyz, mesh.conn, belem, bside, u, s_a=0.8, rowptr=rp, colidx=ci)
        Jo, do = Jo + np.abs(Jx), do + np.abs(dx)
        Jr_b, dr_b = Jx, dx
    else:
        Jr_b, dr_b = 0.0, 0.0
    if kind in ("symmetry", "both", "all"):
        se, ss = (belem, bside) if kind == "symmetry" else _bottom(mesh)
        disc.set_symmetry(se, ss)
        Jx, dx = O.outflow(p, et, mesh.xyz, mesh.conn, se, ss, u, s_a=0.8, rowptr=rp, colidx=ci, kind="symmetry")
        Jo, do = Jo + np.abs(Jx), do + np.abs(dx)
        Jr_b, dr_b = Jr_b + Jx, dr_b + dx
    if kind in ("sliding", "all"):
        we, ws = (belem, bside) if kind == "sliding" else _left(mesh)
        disc.set_sliding_wall(we, ws, 1.3, 0.2)
        Jx, dx = O.outflow(p, et, mesh.xyz, mesh.conn, we, ws, u, s_a=0.8, rowptr=rp, colidx=ci, kind="sliding",
                           k=1.3, lim=0.2)
        Jo, do = Jo + np.abs(Jx), do + np.abs(dx)
        Jr_b, dr_b = Jr_b + Jx, dr_b + dx
    J, d = disc.assemble(torch.from_numpy(u).to(cuda), s_a=0.8, parts=PART_A, jprec=jprec)
    J, d = J.cpu().numpy().astype(np.float64), d.cpu().numpy()
    Jr, dr = O.assemble(p, et, mesh.xyz, mesh.conn, u, s_a=0.8, parts=1, rowptr=rp, colidx=ci)
    Ja, da = O.assemble_abs(p, et, mesh.xyz, mesh.conn, u, s_a=0.8, parts=1, rowptr=rp, colidx=ci)
    assert np.abs(Jo).max() > 0 and np.abs(do).max() > 0
    # per-row magnitude of the outflow terms as their rounding scale
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    rmax = np.zeros(len(rp) - 1)
    np.maximum.at(rmax, rows, np.abs(Jo))
    SJ = Ja + rmax[rows]
    Sd = da + np.abs(do)
    Jw, dw = Jr + Jr_b, dr + dr_b
    floor = PRISM_FLOOR if et == M.PRISM else 0.0
    if jprec == F64:
        tolJ = 1e-12 * np.abs(Jw) + 1e-13 * SJ + floor * np.abs(Jw).max()
    else:
        tolJ = 2.0 ** -17 * SJ + 2.0 ** -20 * np.abs(Jw) + floor * np.abs(Jw).max()
    told = 1e-12 * np.abs(dw) + 1e-13 * Sd + floor * np.abs(dw).max()
    vJ = np.max(np.abs(J - Jw) / (tolJ + 1e-300))
    vd = np.max(np.abs(d - dw) / (told + 1e-300))
    print(f"et={et} jprec={jprec} cfg={cfg} kind={kind}: violation J {vJ:.3f} d {vd:.3f}, outflow sides {len(belem)}")
    assert vJ <= 1.0 and vd <= 1.0


def test_outflow_cleared(cuda):
    import torch

    mesh, belem, bside = _case(M.QUAD)
    u = torch.from_numpy(M.random_state(mesh, 2)).to(cuda)
    disc = NavierStokesFV1("u,v,p", "Inner")
    configure_disc(disc, {"upwind": "full"})
    disc.set_mesh(mesh)
    J0, d0 = disc.assemble(u)
    disc.set_outflow(belem, bside)
    J1, d1 = disc.assemble(u)
    disc.set_outflow([], [])
    J2, d2 = disc.assemble(u)
    assert not torch.equal(d0, d1)
    assert torch.equal(J0, J2) and torch.equal(d0, d2)
    disc.set_symmetry(belem, bside)
    _, d3 = disc.assemble(u)
    disc.set_symmetry([], [])
    J4, d4 = disc.assemble(u)
    assert not torch.equal(d0, d3)
    assert torch.equal(J0, J4) and torch.equal(d0, d4)
    disc.set_sliding_wall(belem, bside, 2.0, 0.1)
    _, d5 = disc.assemble(u)
    disc.set_sliding_wall([], [], 0.0, 0.0)
    J6, d6 = disc.assemble(u)
    assert not torch.equal(d0, d5)
    assert torch.equal(J0, J6) and torch.equal(d0, d6)


def test_boundary_disc_errors(cuda):
    from plugin_navierstokes_amd import NSFV1Error

    mesh, belem, bside = _case(M.QUAD)
    disc = NavierStokesFV1("u,v,p", "Inner")
    configure_disc(disc, {"upwind": "full"})
    disc.set_mesh(mesh)
    for setter in (disc.set_outflow, disc.set_symmetry):
        with pytest.raises(NSFV1Error) as e:
            setter([mesh.ne], [0])
        assert e.value.name == "ARGUMENT"
        with pytest.raises(NSFV1Error):
            setter([0], [4])  # a quadrilateral has sides 0..3
    with pytest.raises(NSFV1Error):
        disc.set_sliding_wall(belem, bside, float("nan"), 0.0)
    # a side whose corner is not on the element: none exists by construction,
    # so a valid list is accepted and an empty one removes the condition
    disc.set_sliding_wall(belem, bside, 1.0, 0.5)
    disc.set_sliding_wall([], [], 0.0, 0.0)
