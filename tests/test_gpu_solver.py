"""GPU Newton loop (Dirichlet rows, GMRES with the FP32/FP64 Jacobian,
FP64 defect) against the CPU fixpoint solve on the oracle."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from plugin_navierstokes_amd import F32, F64, NavierStokesFV1
from plugin_navierstokes_amd import meshes as M
from plugin_navierstokes_amd.solver import BDFStepper, NewtonSolver, cavity_dirichlet
from tests._cpu_newton import newton as cpu_newton

pytestmark = pytest.mark.gpu


def _disc(mesh, upwind, nu):
    disc = NavierStokesFV1("u,v,p" if mesh.dim == 2 else "u,v,w,p", "Inner")
    disc.set_stabilization("fields")
    disc.set_upwind(upwind)
    disc.set_kinematic_viscosity(nu)
    disc.set_mesh(mesh)
    return disc


def test_linear_solve_residual(cuda):
    import torch

    mesh = M.cavity_2d(16, M.QUAD, 0.1, 2)
    disc = _disc(mesh, "full", 0.01)
    dofs, vals = cavity_dirichlet(mesh)
    disc.set_dirichlet(dofs)
    u = torch.from_numpy(M.cavity_lid_state(mesh, 1.0, 0.1 * M.random_state(mesh, 1))).to(cuda)
    for jprec in [F64, F32]:
        J, d = disc.assemble(u, jprec=jprec)
        disc.apply_dirichlet(J, jprec, d, u, torch.as_tensor(vals, device=cuda))
        b = torch.randn(disc.num_dofs(), dtype=torch.float64, device=cuda)
        x = torch.empty_like(b)
        it, rr = disc.linear_solve(J, b, x, jprec=jprec, rtol=1e-8, maxit=2000, restart=40)
        y = torch.empty_like(b)
        disc.spmv(J, x, y, jprec=jprec)
        true = float(torch.linalg.vector_norm(y - b) / torch.linalg.vector_norm(b))
        print(f"jprec={jprec} GMRES its={it} est={rr:.2e} true={true:.2e}")
        assert rr <= 1e-8 and true <= 1e-7


@pytest.mark.parametrize("jprec", [F64, F32])
def test_steady_cavity_matches_cpu(cuda, jprec):
    """Re = 100 cavity, 24x24 jittered quads, full upwind + FIELDS: the GPU
    Newton solution equals the CPU direct-solve Newton solution."""
    import torch

    mesh = M.cavity_2d(24, M.QUAD, 0.1, 5)
    disc = _disc(mesh, "full", 0.01)
    dofs, vals = cavity_dirichlet(mesh)
    nt = NewtonSolver(disc, dofs, vals, jprec=jprec, rtol=1e-10, maxit=80, lin_rtol=1e-6, lin_maxit=2000, restart=40)
    u = torch.zeros(mesh.nv, 3, dtype=torch.float64, device=cuda)
    st = nt.solve(u)
    print(f"GPU Newton jprec={jprec}: its={st.iterations} lin={st.linear_iterations} "
          f"res {st.residuals[0]:.3e} -> {st.residuals[-1]:.3e}")
    assert st.converged
    p = O.Params(dim=2, upwind="full", nu=0.01)
    ucpu, hist = cpu_newton(p, mesh, dofs=np.sort(dofs), vals=vals[np.argsort(dofs)], rtol=1e-12, maxit=80)
    err = np.abs(u.cpu().numpy() - ucpu).max()
    print(f"max |u_gpu - u_cpu| = {err:.3e}")
    assert err < 1e-7


def test_bdf2_steps_match_cpu(cuda):
    """Two BDF steps (BDF1 start, BDF2) of a small 3D hexahedral cavity with
    skewed upwind against the same scheme on the CPU oracle."""
    import torch

    mesh = M.cavity_3d(4, M.HEX)
    nu, dt = 1.0 / 400.0, 0.01
    disc = _disc(mesh, "skewed", nu)
    dofs, vals = cavity_dirichlet(mesh)
    nt = NewtonSolver(disc, dofs, vals, jprec=F64, rtol=1e-11, maxit=60, lin_rtol=1e-8, lin_maxit=3000, restart=50)
    stepper = BDFStepper(nt, dt)
    u0 = M.cavity_lid_state(mesh, 1.0, M.divergence_free_perturbation(mesh, 1e-3, seed=3))
    u_n = torch.from_numpy(u0).to(cuda)
    gpu = []
    for _ in range(2):
        u_n, st = stepper.step(u_n, guess=stepper.extrapolate(u_n))
        assert st.converged
        gpu.append(u_n.cpu().numpy())
    # CPU: same BDF1 / BDF2 residuals with the oracle's M and A parts
    p = O.Params(dim=3, upwind="skewed", nu=nu)
    rp, ci = O.pattern(O.HEX, mesh.nv, mesh.conn)
    order = np.argsort(dofs)
    dd, dv = dofs[order], vals[order]
    prev, cur = None, u0
    for step in range(2):
        if prev is None:
            s_m = 1.0 / dt
            _, cM = O.assemble(p, O.HEX, mesh.xyz, mesh.conn, cur, s_a=0.0, s_m=1.0, parts=2, rowptr=rp, colidx=ci)
            const = -cM / dt
        else:
            s_m = 1.5 / dt
            w = -2.0 * cur + 0.5 * prev
            _, cM = O.assemble(p, O.HEX, mesh.xyz, mesh.conn, w, s_a=0.0, s_m=1.0, parts=2, rowptr=rp, colidx=ci)
            const = cM / dt
        unew, hist = cpu_newton(p, mesh, u0=(cur if prev is None else 2 * cur - prev), dofs=dd, vals=dv, rtol=1e-12,
                                maxit=60, uold=cur, dt=dt, s_m=s_m, const=const)
        prev, cur = cur, unew
        err = np.abs(gpu[step] - cur).max()
        print(f"BDF step {step}: max |u_gpu - u_cpu| = {err:.3e}")
        assert err < 1e-8
