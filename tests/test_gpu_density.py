"""Function-valued density (set_density with CplUserData,
navier_stokes_fv1.cpp:118-121) through the C ABI: the caller's function is
evaluated at the SCVF and SCV integration points the library reports and
the values are imported there (m_imDensitySCVF, m_imDensitySCV,
navier_stokes_fv1.cpp:193-194, 244-245).  Checked against the oracle with
the same imports: every element type, the generic path (full, LPS upwinds)
and the coupled path (positive upwind, FLOW), FP64 and FP32 Jacobians, A + M
+ rhs parts with the time-dependent stabilisation, together with a
function-valued viscosity; a mixed hexahedron + prism mesh; the stabilised
velocity export; a constant function against the constant (bitwise, same
kernel); the refused combinations and clearing."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from plugin_navierstokes_amd import F32, F64, PART_A, PART_M, PART_RHS, NavierStokesFV1, NSFV1Error
from plugin_navierstokes_amd import meshes as M
from tests._parity import FP32_ABS, FP32_REL, FP64_ABS, FP64_REL, configure_disc, oracle_params, small_mesh, worst

pytestmark = pytest.mark.gpu

ETS = [M.TRI, M.QUAD, M.TET, M.HEX, M.PRISM]
NAMES = {M.TRI: "tri", M.QUAD: "quad", M.TET: "tet", M.HEX: "hex", M.PRISM: "prism"}
PARTS = PART_A | PART_M | PART_RHS
# meshes with prisms: FV1 terms that cancel exactly in the oracle can leave
# rounding residues on the GPU; the floor is 1e-15 of the largest entry
PRISM_FLOOR = 1e-15


def RHO(x):
    return 1.0 + 0.6 * x[:, 0] - 0.3 * x[:, 1] ** 2


def NU(x):
    return 0.02 * (1.0 + x[:, 0] + 0.5 * x[:, 1] ** 2)


def _disc(mesh, cfg):
    disc = NavierStokesFV1("u,v,p" if mesh.dim == 2 else "u,v,w,p", "Inner")
    configure_disc(disc, cfg)
    disc.set_mesh(mesh)
    return disc


def _oracle_ips(et, conn, xyz):
    return np.stack([O.geometry(et, xyz[conn[e]])["ipg"] for e in range(conn.shape[0])])


def _group_params(p, et, conn, xyz):
    q = O.Params(**{k: getattr(p, k) for k in p.__dataclass_fields__})
    xo = _oracle_ips(et, conn, xyz)
    ne, dim = conn.shape[0], xyz.shape[1]
    q.rho_ip = RHO(xo.reshape(-1, dim)).reshape(ne, -1)
    q.rho_scv = RHO(xyz[conn].reshape(-1, dim)).reshape(ne, -1)
    q.nu_ip = NU(xo.reshape(-1, dim)).reshape(ne, -1)
    return q


def _viol(got, want, scale, rel, ab, prisms):
    """max |got - want| / (rel |want| + ab scale + floor); <= 1 passes.
    The floor is 1e-15 max|want| on meshes with prisms, else 0."""
    floor = PRISM_FLOOR * np.abs(want).max() if prisms else 0.0
    tol = rel * np.abs(want) + ab * np.asarray(scale) + floor + 1e-300
    return float(np.max(np.abs(np.asarray(got, dtype=np.float64) - want) / tol))


CFGS = [{"upwind": "full"}, {"upwind": "lps"}, {"upwind": "positive"}, {"upwind": "lps", "stab": "flow"}]


@pytest.mark.parametrize("jprec", [F64, F32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("cfg", CFGS, ids=lambda c: "-".join(c.values()))
@pytest.mark.parametrize("et", ETS, ids=lambda e: NAMES[e])
def test_function_valued_density(cuda, et, cfg, jprec):
    import torch

    print(f"case et={NAMES[et]} cfg={cfg} jprec={jprec}", flush=True)
    mesh = small_mesh(et)
    coupled = cfg.get("stab") == "flow" or cfg["upwind"] == "positive"
    disc = _disc(mesh, cfg)
    disc.set_density(RHO)
    disc.set_kinematic_viscosity(NU)
    u, uold = M.random_state(mesh, 15), M.random_state(mesh, 115)
    dt, s = 0.01, (0.7, 1.3, -0.7)
    J, d = disc.assemble(torch.from_numpy(u).to(cuda), torch.from_numpy(uold).to(cuda), dt, s_a=s[0], s_m=s[1],
                         s_f=s[2], parts=PARTS, jprec=jprec)
    assert disc.assembly_path_used() == (4 if coupled else 1)
    J, d = J.cpu().numpy(), d.cpu().numpy()
    p = oracle_params(mesh.dim, cfg)
    p.source = (0.2, -0.1, 0.3)[: mesh.dim]
    disc.set_source(list(p.source))
    J, d = disc.assemble(torch.from_numpy(u).to(cuda), torch.from_numpy(uold).to(cuda), dt, s_a=s[0], s_m=s[1],
                         s_f=s[2], parts=PARTS, jprec=jprec)
    J, d = J.cpu().numpy(), d.cpu().numpy()
    q = _group_params(p, et, mesh.conn, mesh.xyz)
    rowptr, colidx = disc.pattern()
    kw = dict(s_a=s[0], s_m=s[1], rowptr=rowptr, colidx=colidx, parts=PARTS)
    Jr, dr = O.assemble(q, et, mesh.xyz, mesh.conn, u, uold, dt, s_f=s[2], **kw)
    Ja, da = O.assemble_abs(q, et, mesh.xyz, mesh.conn, u, uold, dt, **kw)
    da = da + np.abs(dr)
    # the density import changes the result (the test is not vacuous)
    q0 = _group_params(p, et, mesh.conn, mesh.xyz)
    q0.rho_ip = q0.rho_scv = None
    _, d0 = O.assemble(q0, et, mesh.xyz, mesh.conn, u, uold, dt, s_f=s[2], **kw)
    assert np.max(np.abs(d0 - dr)) > 1e-3 * np.max(np.abs(dr))
    pr = et == M.PRISM
    assert _viol(d, dr, da, FP64_REL, FP64_ABS, pr) <= 1.0, worst(d, dr, da)
    if jprec == F64:
        assert _viol(J, Jr, Ja, FP64_REL, FP64_ABS, pr) <= 1.0, worst(J, Jr, Ja)
    else:
        assert _viol(J, Jr, Ja, FP32_REL, FP32_ABS, pr) <= 1.0


@pytest.mark.parametrize("cfg", [{"upwind": "full"}, {"upwind": "positive"}], ids=["full", "positive"])
def test_function_valued_density_on_a_mixed_mesh(cuda, cfg):
    import torch

    print(f"case mixed hex+prism cfg={cfg}", flush=True)
    mesh = M.mixed_3d(3, "hex_prism")
    disc = _disc(mesh, cfg)
    disc.set_density(RHO)
    disc.set_kinematic_viscosity(NU)
    u, uold = M.random_state(mesh, 21), M.random_state(mesh, 22)
    dt, s = 0.01, (1.0, 1.0, 1.0)
    J, d = disc.assemble(torch.from_numpy(u).to(cuda), torch.from_numpy(uold).to(cuda), dt, s_a=s[0], s_m=s[1],
                         s_f=s[2], parts=PARTS, jprec=F64)
    J, d = J.cpu().numpy(), d.cpu().numpy()
    p = oracle_params(3, cfg)
    rowptr, colidx = disc.pattern()
    Jr = dr = Ja = da = 0.0
    for et, conn in mesh.groups:
        q = _group_params(p, et, conn, mesh.xyz)
        kw = dict(s_a=s[0], s_m=s[1], rowptr=rowptr, colidx=colidx, parts=PARTS)
        a, b = O.assemble(q, et, mesh.xyz, conn, u, uold, dt, s_f=s[2], **kw)
        c, e = O.assemble_abs(q, et, mesh.xyz, conn, u, uold, dt, **kw)
        Jr, dr, Ja, da = Jr + a, dr + b, Ja + c, da + e
    da = da + np.abs(dr)
    assert _viol(d, dr, da, FP64_REL, FP64_ABS, True) <= 1.0, worst(d, dr, da)
    assert _viol(J, Jr, Ja, FP64_REL, FP64_ABS, True) <= 1.0, worst(J, Jr, Ja)


@pytest.mark.parametrize("cfg", [{"upwind": "full"}, {"upwind": "positive"}], ids=["full", "positive"])
@pytest.mark.parametrize("et", [M.QUAD, M.HEX], ids=lambda e: NAMES[e])
def test_stabilised_velocity_export_with_density(cuda, et, cfg):
    import torch

    mesh = small_mesh(et)
    disc = _disc(mesh, cfg)
    disc.set_density(RHO)
    u = M.random_state(mesh, 31)
    out = disc.export(torch.from_numpy(u).to(cuda), "velocity_ip").cpu().numpy()
    q = _group_params(oracle_params(mesh.dim, cfg), et, mesh.conn, mesh.xyz)
    q.nu_ip = None
    ref = O.export(q, et, mesh.xyz, mesh.conn, u, "velocity_ip")
    q.rho_ip = q.rho_scv = None
    ref0 = O.export(q, et, mesh.xyz, mesh.conn, u, "velocity_ip")
    sc = np.abs(ref).max()
    assert np.abs(ref - ref0).max() > 1e-4 * sc
    assert np.abs(out.reshape(ref.shape) - ref).max() <= 1e-11 * sc


@pytest.mark.parametrize("et", [M.TRI, M.HEX], ids=lambda e: NAMES[e])
def test_constant_density_function_equals_the_constant(cuda, et):
    import torch

    mesh = small_mesh(et)
    a = _disc(mesh, {"upwind": "lps", "rho": 1.7, "source": (0.3, -0.2, 0.1)[: mesh.dim]})
    a.set_assembly_path(1)
    b = _disc(mesh, {"upwind": "lps", "source": (0.3, -0.2, 0.1)[: mesh.dim]})
    b.set_density(lambda x: np.full(x.shape[0], 1.7))
    u = torch.from_numpy(M.random_state(mesh, 2)).to(cuda)
    uo = torch.from_numpy(M.random_state(mesh, 3)).to(cuda)
    for jprec in (F64, F32):
        Ja, da = a.assemble(u, uo, 0.01, s_m=1.0, s_f=1.0, parts=PARTS, jprec=jprec)
        Jb, db = b.assemble(u, uo, 0.01, s_m=1.0, s_f=1.0, parts=PARTS, jprec=jprec)
        assert a.assembly_path_used() == b.assembly_path_used() == 1
        assert torch.equal(Ja, Jb) and torch.equal(da, db)


def test_density_refusals_and_clearing(cuda):
    import torch

    mesh = small_mesh(M.HEX)
    u = torch.from_numpy(M.random_state(mesh, 2)).to(cuda)
    disc = _disc(mesh, {"upwind": "full"})
    with pytest.raises(NSFV1Error) as e:  # wrong sizes
        disc.set_density((np.ones((mesh.ne, 11)), np.ones((mesh.ne, 8))))
    assert e.value.name == "ARGUMENT"
    with pytest.raises(NSFV1Error) as e:  # zero density at an SCVF ip
        disc.set_density((np.zeros((mesh.ne, 12)), np.ones((mesh.ne, 8))))
    assert e.value.name == "ARGUMENT"
    disc.set_density(RHO)
    disc.assemble(u, jprec=F32)
    assert disc.assembly_path_used() == 1
    # boundary discs keep a constant density: refused with the import
    belem, bside = M.boundary_sides(mesh, mesh.boundary["lid"])
    disc.set_outflow(belem, bside)
    with pytest.raises(NSFV1Error) as e:
        disc.assemble(u, jprec=F64)
    assert e.value.name == "NOT_IMPLEMENTED"
    disc.set_outflow([], [])
    disc.assemble(u, jprec=F64)
    # a constant clears the import: the default path runs again
    disc.set_density(1.0)
    disc.assemble(u, jprec=F32)
    assert disc.assembly_path_used() != 1
    # a function is evaluated again on a new mesh (other element count)
    disc.set_density(RHO)
    mesh2 = small_mesh(M.HEX, n=6)
    disc.set_mesh(mesh2)
    u2 = torch.from_numpy(M.random_state(mesh2, 2)).to(cuda)
    disc.assemble(u2, jprec=F32)
    assert disc.assembly_path_used() == 1
