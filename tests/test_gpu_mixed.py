"""Mixed-element meshes through the C ABI (nsfv1_set_mesh_mixed): several
element types over one node set, as ugcore's element loop runs the disc
once per type of a subset (navier_stokes_fv1.cpp:136-182, :1515-1551).
Checked against the oracle's sum over the element groups in the same
pattern (quadrilaterals + triangles in both group orders, hexahedra +
prisms, prisms + tetrahedra; four upwinds; FP64 and FP32 Jacobians; A + M +
rhs parts, time-dependent stabilisation); the pattern against the union of
the types' adjacencies; one element type given as two groups against the
single group bitwise (the same sums in the same order), on the generic and
the coupled path; the coupled path (positive upwind, FLOW) on
every mixed mesh; function-valued viscosity and source, group
by group; boundary discs on sides of both element types; data exports
group by group; the skewed-upwind record and replay across groups; a Newton
solve; and the refusals.

Tolerance: the parity rule (|a - b| <= 1e-12 |b| + 1e-13 S in FP64,
2^-17 S + 2^-20 |b| for FP32 Jacobians); on meshes with prisms plus the
prisms' absolute floor of 1e-15 max|b| (an entry that cancels exactly in the
oracle's corner order can be ~1e-19 in the kernel's canonical order)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from plugin_navierstokes_amd import (F32, F64, PART_A, PART_M, PART_RHS, UPWIND_FRESH, UPWIND_RECORD,
                                     UPWIND_REPLAY, NavierStokesFV1, NSFV1Error)
from plugin_navierstokes_amd import meshes as M
from tests._parity import FP32_ABS, FP32_REL, FP64_ABS, FP64_REL, configure_disc, oracle_params

PRISM_FLOOR = 1e-15


def _viol(a, b, scale, rel, ab, floor):
    a, b, scale = np.asarray(a, dtype=np.float64), np.asarray(b), np.asarray(scale)
    tol = rel * np.abs(b) + ab * scale + floor * np.abs(b).max() + 1e-300
    i = int(np.argmax(np.abs(a - b) / tol))
    return float(np.abs(a - b)[i] / tol[i]), f"idx={i} got={a[i]!r} want={b[i]!r} scale={scale[i]!r}"

pytestmark = pytest.mark.gpu

CASES = {
    "quad_tri": lambda: M.mixed_2d(6, 0.1, 1),
    "tri_quad": lambda: M.mixed_2d(5, 0.1, 2, first=M.TRI),
    "hex_prism": lambda: M.mixed_3d(3, "hex_prism", 0.1, 1),
    "prism_tet": lambda: M.mixed_3d(4, "prism_tet", 0.1, 2),
}


def _disc(mm, cfg):
    disc = NavierStokesFV1("u,v,p" if mm.dim == 2 else "u,v,w,p", "Inner")
    configure_disc(disc, cfg)
    disc.set_mesh(mm)
    return disc


@pytest.mark.parametrize("jprec", [F64, F32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("up", ["no", "full", "skewed", "lps"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_mixed_mesh_parity(cuda, name, up, jprec):
    import torch

    print(f"case mixed={name} upwind={up} jprec={jprec}", flush=True)
    mm = CASES[name]()
    dim = mm.dim
    cfg = {"upwind": up, "source": (0.3, -0.2, 0.1)[:dim]}
    disc = _disc(mm, cfg)
    rowptr, colidx = disc.pattern()
    rp, ci = O.pattern_mixed(mm.nv, mm.groups)
    assert np.array_equal(rowptr, rp) and np.array_equal(colidx, ci)
    u = M.random_state(mm, 5)
    uold = M.random_state(mm, 105)
    dt, s = 0.01, (0.7, 1.3, -0.7)
    parts = PART_A | PART_M | PART_RHS
    J, d = disc.assemble(torch.from_numpy(u).to(cuda), torch.from_numpy(uold).to(cuda), dt, s_a=s[0], s_m=s[1],
                         s_f=s[2], parts=parts, jprec=jprec)
    assert disc.assembly_path_used() == 1
    J, d = J.cpu().numpy(), d.cpu().numpy()
    p = oracle_params(dim, cfg)
    kw = dict(uold=uold, dt=dt, s_a=s[0], s_m=s[1], parts=parts)
    Jr, dr = O.assemble_mixed(p, mm.groups, mm.xyz, u, rowptr, colidx, s_f=s[2], **kw)
    Ja, da = O.assemble_abs_mixed(p, mm.groups, mm.xyz, u, rowptr, colidx, **kw)
    da = da + np.abs(dr)
    floor = PRISM_FLOOR if any(et == M.PRISM for et, _ in mm.groups) else 0.0
    v, where = _viol(d, dr, da, FP64_REL, FP64_ABS, floor)
    assert v <= 1.0, where
    if jprec == F64:
        v, where = _viol(J, Jr, Ja, FP64_REL, FP64_ABS, floor)
    else:
        v, where = _viol(J, Jr, Ja, FP32_REL, FP32_ABS, floor)
    assert v <= 1.0, where


COUPLED = [{"upwind": "positive"}, {"upwind": "lps", "stab": "flow"}, {"upwind": "positive", "stab": "flow"}]


def _cid(c):
    return "-".join(f"{k}={v}" for k, v in c.items())


@pytest.mark.parametrize("jprec", [F64, F32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("cfg", COUPLED, ids=_cid)
@pytest.mark.parametrize("name", sorted(CASES))
def test_mixed_mesh_coupled_parity(cuda, name, cfg, jprec):
    """The coupled path (positive upwind, FLOW) group by group."""
    import torch

    print(f"case mixed={name} cfg={cfg} jprec={jprec}", flush=True)
    mm = CASES[name]()
    dim = mm.dim
    cfg = dict(cfg, source=(0.3, -0.2, 0.1)[:dim])
    disc = _disc(mm, cfg)
    rowptr, colidx = disc.pattern()
    u = M.random_state(mm, 6)
    uold = M.random_state(mm, 106)
    dt, s = 0.01, (0.7, 1.3, -0.7)
    parts = PART_A | PART_M | PART_RHS
    J, d = disc.assemble(torch.from_numpy(u).to(cuda), torch.from_numpy(uold).to(cuda), dt, s_a=s[0], s_m=s[1],
                         s_f=s[2], parts=parts, jprec=jprec)
    assert disc.assembly_path_used() == 4
    J, d = J.cpu().numpy(), d.cpu().numpy()
    p = oracle_params(dim, cfg)
    kw = dict(uold=uold, dt=dt, s_a=s[0], s_m=s[1], parts=parts)
    Jr, dr = O.assemble_mixed(p, mm.groups, mm.xyz, u, rowptr, colidx, s_f=s[2], **kw)
    Ja, da = O.assemble_abs_mixed(p, mm.groups, mm.xyz, u, rowptr, colidx, **kw)
    da = da + np.abs(dr)
    floor = PRISM_FLOOR if any(et == M.PRISM for et, _ in mm.groups) else 0.0
    v, where = _viol(d, dr, da, FP64_REL, FP64_ABS, floor)
    assert v <= 1.0, where
    if jprec == F64:
        v, where = _viol(J, Jr, Ja, FP64_REL, FP64_ABS, floor)
    else:
        v, where = _viol(J, Jr, Ja, FP32_REL, FP32_ABS, floor)
    assert v <= 1.0, where


@pytest.mark.parametrize("cfg", COUPLED[:2], ids=_cid)
@pytest.mark.parametrize("et", [M.QUAD, M.TET, M.HEX], ids=["quad", "tet", "hex"])
def test_one_type_as_two_groups_coupled_is_bitwise(cuda, et, cfg):
    import torch

    mesh = M.cavity_2d(6, et, 0.1, 3) if et == M.QUAD else M.cavity_3d(3, et, 0.1, 3)
    print(f"case split groups et={et} cfg={cfg}", flush=True)
    a = NavierStokesFV1("u,v,p" if mesh.dim == 2 else "u,v,w,p", "Inner")
    configure_disc(a, cfg)
    a.set_mesh(mesh)
    b = _disc(M.split_groups(mesh, mesh.ne // 3), cfg)
    u = torch.from_numpy(M.random_state(mesh, 2)).to(cuda)
    uo = torch.from_numpy(M.random_state(mesh, 3)).to(cuda)
    for jprec in (F64, F32):
        Ja, da = a.assemble(u, uo, 0.01, s_m=1.0, parts=PART_A | PART_M, jprec=jprec)
        Jb, db = b.assemble(u, uo, 0.01, s_m=1.0, parts=PART_A | PART_M, jprec=jprec)
        assert a.assembly_path_used() == b.assembly_path_used() == 4
        assert torch.equal(Ja, Jb) and torch.equal(da, db)


@pytest.mark.parametrize("et", [M.TRI, M.TET, M.HEX, M.PRISM], ids=["tri", "tet", "hex", "prism"])
def test_one_type_as_two_groups_is_bitwise_the_single_group(cuda, et):
    import torch

    mesh = M.cavity_2d(6, et, 0.1, 3) if et == M.TRI else M.cavity_3d(3, et, 0.1, 3)
    print(f"case split groups et={et} ne={mesh.ne}", flush=True)
    cfg = {"upwind": "lps"}
    a = NavierStokesFV1("u,v,p" if mesh.dim == 2 else "u,v,w,p", "Inner")
    configure_disc(a, cfg)
    a.set_mesh(mesh)
    a.set_assembly_path(1)
    b = _disc(M.split_groups(mesh, mesh.ne // 3), cfg)
    u = torch.from_numpy(M.random_state(mesh, 2)).to(cuda)
    uo = torch.from_numpy(M.random_state(mesh, 3)).to(cuda)
    for jprec in (F64, F32):
        Ja, da = a.assemble(u, uo, 0.01, s_m=1.0, parts=PART_A | PART_M, jprec=jprec)
        Jb, db = b.assemble(u, uo, 0.01, s_m=1.0, parts=PART_A | PART_M, jprec=jprec)
        assert a.assembly_path_used() == b.assembly_path_used() == 1
        assert torch.equal(Ja, Jb) and torch.equal(da, db)
        # defect only (no Jacobian) through the same groups
        _, dd = b.assemble(u, uo, 0.01, s_m=1.0, parts=PART_A | PART_M, jprec=jprec, want_J=False)
        assert torch.equal(dd, db)


def test_skewed_record_and_replay_across_groups(cuda):
    import torch

    mm = CASES["hex_prism"]()
    disc = _disc(mm, {"upwind": "skewed"})
    u = torch.from_numpy(M.random_state(mm, 7)).to(cuda)
    u2 = torch.from_numpy(M.random_state(mm, 8)).to(cuda)
    disc.set_upwind_decisions(UPWIND_RECORD)
    J1, d1 = disc.assemble(u, jprec=F64)
    disc.set_upwind_decisions(UPWIND_REPLAY)
    J2, d2 = disc.assemble(u, jprec=F64)
    assert torch.equal(J1, J2) and torch.equal(d1, d2)
    # replayed decisions at another state differ from fresh ones there
    J3, _ = disc.assemble(u2, jprec=F64)
    disc.set_upwind_decisions(UPWIND_FRESH)
    J4, _ = disc.assemble(u2, jprec=F64)
    assert not torch.equal(J3, J4)


def test_newton_on_a_mixed_mesh(cuda):
    """A Re = 20 cavity of quadrilaterals and triangles solved by the device
    Newton/GMRES loop: the converged state's oracle defect on the free rows
    is below 1e-8 of the initial one."""
    import torch

    from plugin_navierstokes_amd.solver import NewtonSolver, cavity_dirichlet

    mm = M.mixed_2d(8, 0.1, 4)
    cfg = {"upwind": "full", "nu": 0.05}
    disc = _disc(mm, cfg)
    dofs, vals = cavity_dirichlet(mm)
    o = np.argsort(dofs)
    dofs, vals = dofs[o], vals[o]
    nt = NewtonSolver(disc, dofs, vals, jprec=F64, rtol=1e-10, maxit=40, lin_rtol=1e-6, lin_maxit=400, restart=30)
    u = torch.zeros(mm.nv, 3, dtype=torch.float64, device=cuda)
    st = nt.solve(u)
    print(f"mixed Newton: its={st.iterations} res {st.residuals[0]:.3e} -> {st.residuals[-1]:.3e}", flush=True)
    assert st.converged
    rowptr, colidx = disc.pattern()
    p = oracle_params(2, cfg)
    free = np.ones(mm.nv * 3, bool)
    free[dofs] = False
    u0 = np.zeros((mm.nv, 3))
    u0.reshape(-1)[dofs] = vals
    _, d0 = O.assemble_mixed(p, mm.groups, mm.xyz, u0, rowptr, colidx)
    _, d1 = O.assemble_mixed(p, mm.groups, mm.xyz, u.cpu().numpy(), rowptr, colidx)
    assert np.linalg.norm(d1[free]) <= 1e-8 * np.linalg.norm(d0[free])


def _nu(x):
    return 0.02 * (1.0 + x[:, 0] + 0.5 * x[:, 1] ** 2)


def _f(x):
    cols = [np.sin(3.0 * x[:, 0]), np.cos(2.0 * x[:, 1]), x[:, 0] * x[:, 1] - 0.2]
    return np.stack(cols[: x.shape[1]], 1)


def _coupled(cfg):
    return cfg["upwind"] == "positive" or cfg.get("stab") == "flow"


IMPORT_CASES = [(n, c) for n in sorted(CASES)
                for c in ({"upwind": "full"}, {"upwind": "lps"}, {"upwind": "positive"}, {"upwind": "lps", "stab": "flow"})]


@pytest.mark.parametrize("name,cfg", IMPORT_CASES, ids=[f"{n}-{_cid(c)}" for n, c in IMPORT_CASES])
def test_mixed_mesh_function_imports(cuda, name, cfg):
    """Function-valued viscosity and source on a mixed mesh: positions and
    imports run through the groups in order; each group's elements take
    their own slice (generic and coupled path)."""
    import torch

    mm = CASES[name]()
    coupled = _coupled(cfg)
    print(f"case mixed={name} cfg={cfg} imports", flush=True)
    dim = mm.dim
    disc = _disc(mm, cfg)
    disc.set_kinematic_viscosity(_nu)
    disc.set_source(_f)
    xf = disc.ip_positions("scvf")
    xo = np.concatenate([np.stack([O.geometry(et, mm.xyz[c])["ipg"] for c in conn]).reshape(-1, dim)
                         for et, conn in mm.groups])
    assert xf.shape == xo.shape and np.max(np.abs(xf - xo)) <= 1e-14
    xc = disc.ip_positions("scv")
    assert np.array_equal(xc, np.concatenate([mm.xyz[conn].reshape(-1, dim) for _, conn in mm.groups]))
    u = M.random_state(mm, 5)
    uold = M.random_state(mm, 105)
    dt, s = 0.01, (0.7, 1.3, -0.7)
    parts = PART_A | PART_M | PART_RHS
    J, d = disc.assemble(torch.from_numpy(u).to(cuda), torch.from_numpy(uold).to(cuda), dt, s_a=s[0], s_m=s[1],
                         s_f=s[2], parts=parts, jprec=F64)
    assert disc.assembly_path_used() == (4 if coupled else 1)
    J, d = J.cpu().numpy(), d.cpu().numpy()
    rowptr, colidx = disc.pattern()
    Jr = dr = Ja = da = 0.0
    for et, conn in mm.groups:
        p = oracle_params(dim, cfg)
        p.source = (0.0,) * dim
        ipg = np.stack([O.geometry(et, mm.xyz[c])["ipg"] for c in conn])
        p.nu_ip = _nu(ipg.reshape(-1, dim)).reshape(len(conn), -1)
        p.f_scv = _f(mm.xyz[conn].reshape(-1, dim)).reshape(len(conn), -1, dim)
        p.f_scvf = _f(ipg.reshape(-1, dim)).reshape(len(conn), -1, dim)
        kw = dict(s_a=s[0], s_m=s[1], parts=parts, rowptr=rowptr, colidx=colidx)
        Jg, dg = O.assemble(p, et, mm.xyz, conn, u, uold, dt, s_f=s[2], **kw)
        Jag, dag = O.assemble_abs(p, et, mm.xyz, conn, u, uold, dt, **kw)
        Jr, dr, Ja, da = Jr + Jg, dr + dg, Ja + Jag, da + dag
    da = da + np.abs(dr)
    floor = PRISM_FLOOR if any(et == M.PRISM for et, _ in mm.groups) else 0.0
    for a, b, sc in ((d, dr, da), (J, Jr, Ja)):
        v, where = _viol(a, b, sc, FP64_REL, FP64_ABS, floor)
        assert v <= 1.0, where


def _split(mm, belem, bside):
    """Global boundary (element, side) lists as per-group local lists."""
    out, e0 = [], 0
    for et, conn in mm.groups:
        m = (belem >= e0) & (belem < e0 + conn.shape[0])
        out.append((et, conn, (belem[m] - e0).astype(np.int32), bside[m].astype(np.int32)))
        e0 += conn.shape[0]
    return out


@pytest.mark.parametrize("jprec", [F64, F32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("kind", ["outflow", "symmetry", "sliding"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_mixed_mesh_boundary_discs(cuda, name, kind, jprec):
    """Boundary discs on sides of both element types (the y = max side meets
    every group): each group's boundary faces are added after the element
    assembly, group by group, against the oracle's discs per group."""
    import torch

    print(f"case mixed={name} boundary={kind} jprec={jprec}", flush=True)
    mm = CASES[name]()
    dim = mm.dim
    cfg = {"upwind": "full"}
    disc = _disc(mm, cfg)
    top = np.nonzero(np.isclose(mm.xyz[:, 1], mm.xyz[:, 1].max()))[0]  # y = max meets every group
    belem, bside = M.boundary_sides(mm, top)
    parts_g = _split(mm, belem, bside)
    assert sum(len(b) > 0 for _, _, b, _ in parts_g) == 2  # both groups have sides on it
    if kind == "outflow":
        disc.set_outflow(belem, bside)
    elif kind == "symmetry":
        disc.set_symmetry(belem, bside)
    else:
        disc.set_sliding_wall(belem, bside, 1.3, 0.2)
    u = M.random_state(mm, 21)
    J, d = disc.assemble(torch.from_numpy(u).to(cuda), s_a=0.8, parts=PART_A, jprec=jprec)
    J, d = J.cpu().numpy().astype(np.float64), d.cpu().numpy()
    rp, ci = disc.pattern()
    p = oracle_params(dim, cfg)
    Jr, dr = O.assemble_mixed(p, mm.groups, mm.xyz, u, rp, ci, s_a=0.8, parts=1)
    Ja, da = O.assemble_abs_mixed(p, mm.groups, mm.xyz, u, rp, ci, s_a=0.8, parts=1)
    Jo = do = 0.0
    for et, conn, be, bs in parts_g:
        if len(be) == 0:
            continue
        kw = dict(kind=kind, k=1.3, lim=0.2) if kind == "sliding" else dict(kind=kind)
        Jx, dx = O.outflow(p, et, mm.xyz, conn, be, bs, u, s_a=0.8, rowptr=rp, colidx=ci, **kw)
        Jr, dr = Jr + Jx, dr + dx
        Jo, do = Jo + np.abs(Jx), do + np.abs(dx)
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    rmax = np.zeros(len(rp) - 1)
    np.maximum.at(rmax, rows, np.abs(Jo))
    floor = PRISM_FLOOR if any(et == M.PRISM for et, _ in mm.groups) else 0.0
    SJ = Ja + rmax[rows]
    if jprec == F64:
        v, where = _viol(J, Jr, SJ, FP64_REL, FP64_ABS, floor)
    else:
        v, where = _viol(J, Jr, SJ, FP32_REL, FP32_ABS, floor)
    assert v <= 1.0, where
    v, where = _viol(d, dr, da + np.abs(do), FP64_REL, FP64_ABS, floor)
    assert v <= 1.0, where


@pytest.mark.parametrize("cfg", [{"upwind": "full"}, {"upwind": "positive", "stab": "flow"}], ids=_cid)
@pytest.mark.parametrize("at", ["scvf", "scv"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_mixed_mesh_exports(cuda, name, at, cfg):
    """Data exports on a mixed mesh: each group's [ne][points][components]
    block in group order, against the oracle's exports per group (the
    stabilised ip velocity through the generic or the coupled path)."""
    import torch

    print(f"case mixed={name} exports at {at} cfg={cfg}", flush=True)
    mm = CASES[name]()
    dim = mm.dim
    disc = _disc(mm, cfg)
    u = M.random_state(mm, 31)
    uold = M.random_state(mm, 32)
    ut, uot = torch.from_numpy(u).to(cuda), torch.from_numpy(uold).to(cuda)
    p = oracle_params(dim, cfg)
    for what in ("velocity", "velocity_grad", "velocity_ip", "pressure", "pressure_grad"):
        got = disc.export(ut, what, at, u_old=uot, dt=0.02).cpu().numpy()
        ncomp = dim * dim if what == "velocity_grad" else (1 if what == "pressure" else dim)
        want = np.concatenate([O.export(p, et, mm.xyz, conn, u, what, at, None, uold, 0.02).reshape(-1, ncomp)
                               for et, conn in mm.groups])
        assert got.shape == want.shape, (what, got.shape, want.shape)
        tol = 1e-12 * np.abs(want) + 1e-12 * np.abs(want).max()
        assert np.max(np.abs(got - want) / tol) <= 1.0, what


def test_refusals(cuda):
    import torch

    mm = CASES["hex_prism"]()
    disc = _disc(mm, {"upwind": "full"})
    u = torch.from_numpy(M.random_state(mm, 2)).to(cuda)
    disc.assemble(u, jprec=F32)
    # local points belong to one reference element
    with pytest.raises(NSFV1Error) as e:
        disc.export(u, "velocity", "local", np.full((2, 3), 0.2))
    assert e.value.name == "NOT_IMPLEMENTED"
    # a pyramid group is refused like a pyramid mesh
    with pytest.raises(NSFV1Error) as e:
        disc.set_mesh(M.MixedMesh([(M.HEX, mm.groups[0][1]), (5, np.zeros((1, 5), np.int32))], mm.xyz, mm.boundary))
    assert e.value.name == "NOT_IMPLEMENTED"
    # a group of the other dimension
    with pytest.raises(NSFV1Error) as e:
        disc.set_mesh(M.MixedMesh([(M.HEX, mm.groups[0][1]), (M.TRI, np.zeros((1, 3), np.int32))], mm.xyz,
                                  mm.boundary))
    assert e.value.name == "ARGUMENT"
    # the handle still works after a refused mesh: a single-type mesh
    mesh = M.cavity_3d(2, M.HEX, 0.1, 1)
    disc.set_mesh(mesh)
    disc.set_kinematic_viscosity(0.05)
    disc.assemble(torch.from_numpy(M.random_state(mesh, 2)).to(cuda), jprec=F32)
