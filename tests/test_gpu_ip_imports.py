"""Function-valued kinematic viscosity and source (the CplUserData overloads
of set_kinematic_viscosity / set_source, register_navier_stokes.cpp:109-118)
through the C ABI: the library evaluates nothing itself; the caller's
function is evaluated at the integration points the library reports
(nsfv1_ip_positions) and the values are imported there
(navier_stokes_fv1.cpp:188-247).  Checked against the oracle with the same
imports (every element type, three upwinds, FP64 and FP32 Jacobians, A + M
+ rhs parts, time-dependent stabilisation); the positions against the
oracle's FV1 geometry; constant functions against the constants (bitwise,
same kernel); the coupled path (positive upwind, FLOW) with the same
imports; multigrid with the functions on every level; the refused
combinations and clearing."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from plugin_navierstokes_amd import F32, F64, PART_A, PART_M, PART_RHS, NavierStokesFV1, NSFV1Error
from plugin_navierstokes_amd import meshes as M
from tests._parity import configure_disc, fp32_violation, fp64_violation, oracle_params, small_mesh, worst

pytestmark = pytest.mark.gpu

ETS = [M.TRI, M.QUAD, M.TET, M.HEX]
NAMES = {M.TRI: "tri", M.QUAD: "quad", M.TET: "tet", M.HEX: "hex"}


def NU(x):
    return 0.02 * (1.0 + x[:, 0] + 0.5 * x[:, 1] ** 2)


def F(x):
    cols = [np.sin(3.0 * x[:, 0]), np.cos(2.0 * x[:, 1]), x[:, 0] * x[:, 1] - 0.2]
    return np.stack(cols[: x.shape[1]], 1)


def _disc(mesh, cfg):
    disc = NavierStokesFV1("u,v,p" if mesh.dim == 2 else "u,v,w,p", "Inner")
    configure_disc(disc, cfg)
    disc.set_mesh(mesh)
    return disc


def _oracle_ips(et, mesh):
    return np.stack([O.geometry(et, mesh.xyz[mesh.conn[e]])["ipg"] for e in range(mesh.ne)])


CFGS = [{"upwind": "full"}, {"upwind": "skewed"}, {"upwind": "lps"},
        # the coupled path: the matrix case of the positive upwind, FLOW
        {"upwind": "positive"}, {"upwind": "lps", "stab": "flow"}, {"upwind": "positive", "stab": "flow"}]


def _cid(c):
    return "-".join(f"{k}={v}" for k, v in c.items())


@pytest.mark.parametrize("jprec", [F64, F32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("cfg", CFGS, ids=_cid)
@pytest.mark.parametrize("et", ETS, ids=lambda e: NAMES[e])
def test_function_valued_viscosity_and_source(cuda, et, cfg, jprec):
    import torch

    print(f"case et={NAMES[et]} cfg={cfg} jprec={jprec}", flush=True)
    mesh = small_mesh(et)
    dim, ne = mesh.dim, mesh.ne
    coupled = cfg.get("stab") == "flow" or cfg["upwind"] == "positive"
    disc = _disc(mesh, cfg)
    disc.set_kinematic_viscosity(NU)
    disc.set_source(F)
    xf = disc.ip_positions("scvf")
    xo = _oracle_ips(et, mesh)
    assert np.max(np.abs(xf - xo)) <= 1e-14
    xc = disc.ip_positions("scv")
    assert np.array_equal(xc, mesh.xyz[mesh.conn])
    u = M.random_state(mesh, 5)
    uold = M.random_state(mesh, 105)
    dt, s = 0.01, (0.7, 1.3, -0.7)
    parts = PART_A | PART_M | PART_RHS
    J, d = disc.assemble(torch.from_numpy(u).to(cuda), torch.from_numpy(uold).to(cuda), dt, s_a=s[0], s_m=s[1],
                         s_f=s[2], parts=parts, jprec=jprec)
    assert disc.assembly_path_used() == (4 if coupled else 1)
    J, d = J.cpu().numpy(), d.cpu().numpy()
    p = oracle_params(dim, cfg)
    p.source = (0.0,) * dim
    p.nu_ip = NU(xo.reshape(-1, dim)).reshape(ne, -1)
    p.f_scv = F(mesh.xyz[mesh.conn].reshape(-1, dim)).reshape(ne, -1, dim)
    p.f_scvf = F(xo.reshape(-1, dim)).reshape(ne, -1, dim)
    rowptr, colidx = disc.pattern()
    Jr, dr = O.assemble(p, et, mesh.xyz, mesh.conn, u, uold, dt, s_a=s[0], s_m=s[1], s_f=s[2], parts=parts,
                        rowptr=rowptr, colidx=colidx)
    Ja, da = O.assemble_abs(p, et, mesh.xyz, mesh.conn, u, uold, dt, s_a=s[0], s_m=s[1], parts=parts,
                            rowptr=rowptr, colidx=colidx)
    da = da + np.abs(dr)
    # the imports differ from the scalar case (the test is not vacuous)
    q = oracle_params(dim, cfg)
    _, d0 = O.assemble(q, et, mesh.xyz, mesh.conn, u, uold, dt, s_a=s[0], s_m=s[1], s_f=s[2], parts=parts,
                       rowptr=rowptr, colidx=colidx)
    assert np.max(np.abs(d0 - dr)) > 1e-3 * np.max(np.abs(dr))
    assert fp64_violation(d, dr, da) <= 1.0, worst(d, dr, da)
    if jprec == F64:
        assert fp64_violation(J, Jr, Ja) <= 1.0, worst(J, Jr, Ja)
    else:
        assert fp32_violation(J, Jr, Ja) <= 1.0


@pytest.mark.parametrize("et", ETS, ids=lambda e: NAMES[e])
def test_constant_functions_equal_the_constants(cuda, et):
    import torch

    mesh = small_mesh(et)
    dim = mesh.dim
    f = (0.3, -0.2, 0.1)[:dim]
    a = _disc(mesh, {"upwind": "lps", "nu": 0.03, "source": f})
    a.set_assembly_path(1)
    b = _disc(mesh, {"upwind": "lps"})
    b.set_kinematic_viscosity(lambda x: np.full(x.shape[0], 0.03))
    b.set_source(lambda x: np.tile(f, (x.shape[0], 1)))
    u = torch.from_numpy(M.random_state(mesh, 2)).to(cuda)
    uo = torch.from_numpy(M.random_state(mesh, 3)).to(cuda)
    for jprec in (F64, F32):
        Ja, da = a.assemble(u, uo, 0.01, s_m=1.0, s_f=1.0, parts=PART_A | PART_M | PART_RHS, jprec=jprec)
        Jb, db = b.assemble(u, uo, 0.01, s_m=1.0, s_f=1.0, parts=PART_A | PART_M | PART_RHS, jprec=jprec)
        assert a.assembly_path_used() == b.assembly_path_used() == 1
        assert torch.equal(Ja, Jb) and torch.equal(da, db)


def test_refusals_and_clearing(cuda):
    import torch

    mesh = small_mesh(M.HEX)
    u = torch.from_numpy(M.random_state(mesh, 2)).to(cuda)
    disc = _disc(mesh, {"upwind": "full"})
    nip = 12
    # wrong sizes
    with pytest.raises(NSFV1Error) as e:
        disc.set_kinematic_viscosity(np.ones((mesh.ne, nip - 1)))
    assert e.value.name == "ARGUMENT"
    disc.set_kinematic_viscosity(NU)
    disc.set_source(F)
    disc.assemble(u, jprec=F32)
    assert disc.assembly_path_used() == 1
    # FLOW and the positive upwind take the imports on the coupled path
    disc.set_stabilization("flow")
    disc.assemble(u, jprec=F64)
    assert disc.assembly_path_used() == 4
    disc.set_stabilization("fields")
    disc.set_upwind("positive")
    disc.assemble(u, jprec=F64)
    assert disc.assembly_path_used() == 4
    disc.set_upwind("full")
    # boundary discs take the function at their faces (the facade evaluates
    # it there); without those values the assembly refuses
    belem, bside = M.boundary_sides(mesh, mesh.boundary["lid"])
    assert len(belem) > 0
    disc.set_outflow(belem, bside)
    disc.assemble(u, jprec=F64)
    disc.set_boundary_viscosity(0, None)
    with pytest.raises(NSFV1Error) as e:
        disc.assemble(u, jprec=F64)
    assert e.value.name == "ARGUMENT"
    disc.set_outflow([], [])
    disc.assemble(u, jprec=F64)
    # constants clear the imports: the default path runs again
    disc.set_kinematic_viscosity(0.05)
    disc.set_source([0.1, 0.0, 0.0])
    disc.assemble(u, jprec=F32)
    assert disc.assembly_path_used() != 1
    # a function is evaluated again on a new mesh (other element count)
    disc.set_kinematic_viscosity(NU)
    mesh2 = small_mesh(M.HEX, n=6)
    disc.set_mesh(mesh2)
    u2 = torch.from_numpy(M.random_state(mesh2, 2)).to(cuda)
    disc.assemble(u2, jprec=F32)
    assert disc.assembly_path_used() == 1


def test_multigrid_with_function_imports(cuda):
    """Newton with the multigrid preconditioner on a cavity with a
    function-valued viscosity and source: the facade evaluates the functions
    on every coarse level too.  The converged state equals the block-Jacobi
    one (the same discrete system) and multigrid needs fewer Krylov
    iterations; an array import given on the fine level only is refused by
    the multigrid set-up."""
    import torch

    from plugin_navierstokes_amd.solver import NewtonSolver, build_multigrid, cavity_dirichlet

    def dirichlet(m):
        dofs, vals = cavity_dirichlet(m)
        o = np.argsort(dofs)
        return dofs[o], vals[o]

    mesh = M.cavity_2d(16, M.QUAD, 0.1, 3)
    dofs, vals = dirichlet(mesh)
    out = {}
    for mg in (False, True):
        print(f"case multigrid={mg}", flush=True)
        disc = _disc(mesh, {"upwind": "full"})
        disc.set_kinematic_viscosity(NU)
        disc.set_source(F)
        if mg:
            build_multigrid(disc, mesh, dirichlet, max_coarse_dofs=300)
        nt = NewtonSolver(disc, dofs, vals, jprec=F64, rtol=1e-11, maxit=40, lin_rtol=1e-6, lin_maxit=600,
                          restart=30, multigrid=mg)
        u = torch.zeros(mesh.nv, 3, dtype=torch.float64, device=cuda)
        st = nt.solve(u)
        print(f"multigrid={mg}: newton {st.iterations}, krylov {st.linear_iterations}", flush=True)
        assert st.converged
        out[mg] = (u.cpu().numpy(), st.linear_iterations)
    scale = np.abs(out[False][0]).max()
    assert np.abs(out[True][0] - out[False][0]).max() <= 1e-7 * scale
    assert out[True][1] < out[False][1]
    # an array import is not evaluated on the coarse levels: refused
    disc = _disc(mesh, {"upwind": "full"})
    build_multigrid(disc, mesh, dirichlet, max_coarse_dofs=300)
    disc.set_kinematic_viscosity(np.full((mesh.ne, 4), 0.02))
    u = torch.zeros(mesh.nv, 3, dtype=torch.float64, device=cuda)
    with pytest.raises(NSFV1Error) as e:
        disc.mg_setup(F64, u)
    assert e.value.name == "ARGUMENT"
    # back to a constant: the set-up runs
    disc.set_kinematic_viscosity(0.02)
    disc.mg_setup(F64, u)
    # a function on the fine level reaches the coarse levels; a constant
    # afterwards drops their imports again at the next set-up
    disc.set_kinematic_viscosity(NU)
    disc.mg_setup(F64, u)
    disc.set_kinematic_viscosity(0.02)
    disc.mg_setup(F64, u)


@pytest.mark.parametrize("jprec", [F64, F32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("kind", ["outflow", "symmetry"])
@pytest.mark.parametrize("et", ETS + [M.PRISM], ids=lambda e: NAMES.get(e, "prism"))
def test_function_viscosity_on_boundary_discs(cuda, et, kind, jprec):
    """The viscosity function at the outflow / symmetry faces
    (no_normal_stress_outflow_fv1.cpp:188-426, symmetric_boundary_fv1.cpp:
    224-386 evaluate the master's viscosity there): the face ips against
    the oracle's boundary faces, J and d against the oracle given the same
    values at the element ips and at the faces."""
    import torch

    print(f"case et={et} boundary={kind} jprec={jprec}", flush=True)
    mesh = small_mesh(et)
    dim = mesh.dim
    cfg = {"upwind": "full"}
    disc = _disc(mesh, cfg)
    disc.set_kinematic_viscosity(NU)
    x = mesh.xyz[:, 0]
    belem, bside = M.boundary_sides(mesh, np.nonzero(np.isclose(x, x.max()))[0])
    (disc.set_outflow if kind == "outflow" else disc.set_symmetry)(belem, bside)
    pos = disc.bnd_ip_positions(0 if kind == "outflow" else 1)
    nu_bf = np.zeros((len(belem), 4))
    for b, (e, sd) in enumerate(zip(belem, bside)):
        ipg = O.bnd_faces(et, mesh.xyz[mesh.conn[e]], int(sd))["ipg"]
        assert np.max(np.abs(pos[b, : len(ipg)] - ipg)) <= 1e-15
        nu_bf[b, : len(ipg)] = NU(ipg)
    u = M.random_state(mesh, 23)
    J, d = disc.assemble(torch.from_numpy(u).to(cuda), s_a=0.8, jprec=jprec)
    J, d = J.cpu().numpy().astype(np.float64), d.cpu().numpy()
    rp, ci = disc.pattern()
    p = oracle_params(dim, cfg)
    xo = _oracle_ips(et, mesh)
    p.nu_ip = NU(xo.reshape(-1, dim)).reshape(mesh.ne, -1)
    Jr, dr = O.assemble(p, et, mesh.xyz, mesh.conn, u, s_a=0.8, rowptr=rp, colidx=ci)
    Ja, da = O.assemble_abs(p, et, mesh.xyz, mesh.conn, u, s_a=0.8, rowptr=rp, colidx=ci)
    p.nu_bf = nu_bf
    Jb, db = O.outflow(p, et, mesh.xyz, mesh.conn, belem, bside, u, s_a=0.8, rowptr=rp, colidx=ci, kind=kind)
    # the constant viscosity gives another answer, far above rounding (the
    # test is not vacuous; the symmetry disc's pressure part does not depend
    # on it)
    q = oracle_params(dim, cfg)
    Jc, _ = O.outflow(q, et, mesh.xyz, mesh.conn, belem, bside, u, s_a=0.8, rowptr=rp, colidx=ci, kind=kind)
    assert np.max(np.abs(Jc - Jb)) > 1e-6 * np.max(np.abs(Jb))
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    rmax = np.zeros(len(rp) - 1)
    np.maximum.at(rmax, rows, np.abs(Jb))
    floor = 1e-15 if et == M.PRISM else 0.0
    Jw, dw = Jr + Jb, dr + db
    SJ = Ja + rmax[rows]
    Sd = da + np.abs(db)
    if jprec == F64:
        tolJ = 1e-12 * np.abs(Jw) + 1e-13 * SJ + floor * np.abs(Jw).max()
    else:
        tolJ = 2.0 ** -17 * SJ + 2.0 ** -20 * np.abs(Jw) + floor * np.abs(Jw).max()
    told = 1e-12 * np.abs(dw) + 1e-13 * Sd + floor * np.abs(dw).max()
    assert np.max(np.abs(J - Jw) / (tolJ + 1e-300)) <= 1.0
    assert np.max(np.abs(d - dw) / (told + 1e-300)) <= 1.0
