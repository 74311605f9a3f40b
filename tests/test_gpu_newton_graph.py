"""The device Newton step (nsfv1_newton_step), its HIP graph
(nsfv1_newton_graph_capture / _replay) and the C++ launch-shape autotuner
(nsfv1_autotune), through the C ABI, on a batch of 2D cavities:
- every graph replay equals the eager step bitwise (u and ||d||^2), with
  eager work and synchronisations between the replays;
- one device step equals assembly + Dirichlet rows + the host-driven GMRES
  (nsfv1_linear_solve, same restart and iteration count) + the update, to
  rounding (the two GMRES differ only in where the Givens rotations run and
  in dividing by ||w|| against multiplying by its reciprocal);
- the step, replayed, drives the batch's defect down by Newton iterations;
- the autotuner keeps a shape of its table, shapes of one path give bitwise
  equal results, shapes a mesh cannot use are skipped;
- misuse is refused (null-stream capture, replay without capture, Krylov
  sizes out of range)."""
from __future__ import annotations

import numpy as np
import pytest

from plugin_navierstokes_amd import F32, F64, PART_A, NSFV1Error, NavierStokesFV1
from plugin_navierstokes_amd import meshes as M
from plugin_navierstokes_amd.batch import BatchedCavities, draw_instances

pytestmark = pytest.mark.gpu


def _batch(jprec=F32, upwind="full"):
    return BatchedCavities(draw_instances(6, seed=3, side=(8, 16), re=(10.0, 100.0)), upwind=upwind, jprec=jprec)


def test_graph_replays_equal_eager_steps(cuda):
    import torch

    b = _batch()
    u0 = b.state(seed=2)
    ue = u0.clone()
    r = torch.zeros(1, dtype=torch.float64, device=cuda)
    eager_u, eager_r = [], []
    for _ in range(4):
        b.newton_step(ue, krylov=20, res2=r)
        torch.cuda.synchronize()
        eager_u.append(ue.clone())
        eager_r.append(r.clone())
    ug = u0.clone()
    rg = torch.zeros(1, dtype=torch.float64, device=cuda)
    b.capture_newton(ug, krylov=20, res2=rg)
    assert torch.equal(ug, u0), "capturing must not run the step"
    for k in range(4):
        b.replay_newton()
        torch.cuda.synchronize()
        junk = torch.randn(1 << 16, device=cuda).sum()  # eager work between the replays
        torch.cuda.synchronize()
        assert torch.isfinite(junk)
        assert torch.equal(ug, eager_u[k]) and torch.equal(rg, eager_r[k]), f"replay {k}"
    b.disc.sync_errors()


@pytest.mark.parametrize("jprec", [F32, F64])
def test_device_step_matches_host_gmres_step(cuda, jprec):
    import torch

    b = _batch(jprec)
    u = b.state(seed=5)
    m = 25
    # host-driven: assembly, Dirichlet rows, GMRES(m) for exactly m iterations, update
    uh = u.clone()
    J, d = b.disc.assemble(uh, jprec=jprec)
    b.disc.apply_dirichlet(J, jprec, d, uh, b.g)
    delta = torch.empty_like(d)
    its, _ = b.disc.linear_solve(J, d, delta, jprec=jprec, rtol=1e-300, maxit=m, restart=m)
    assert its == m
    uh -= delta.view_as(uh)
    # device step
    ud = u.clone()
    r = torch.zeros(1, dtype=torch.float64, device=cuda)
    b.disc.newton_step(ud, g=b.g, krylov=m, res2=r, jprec=jprec)
    torch.cuda.synchronize()
    scale = float(delta.abs().max())
    err = float((ud - uh).abs().max())
    print(f"jprec={jprec}: max |u_dev - u_host| = {err:.3e} of max |delta| = {scale:.3e}")
    assert err <= 1e-10 * scale
    assert abs(float(r) - float(torch.dot(d, d))) <= 1e-12 * float(torch.dot(d, d))


def test_replayed_steps_reduce_the_defect(cuda):
    """Central upwind (continuous in u; the full upwind's corner switches stall
    a fixpoint iteration near 1e-3 relative).  The FV1 Jacobian omits the
    derivatives of the stabilisation weights, so the iteration contracts
    linearly: 12 steps take the defect below 1e-4 of its start."""
    import torch

    b = _batch(upwind="no")
    u = b.state(seed=7)
    r = torch.zeros(1, dtype=torch.float64, device=cuda)
    b.capture_newton(u, krylov=30, res2=r)
    norms = []
    for _ in range(12):
        b.replay_newton()
        torch.cuda.synchronize()
        norms.append(float(r) ** 0.5)
    print("defect norms:", " ".join(f"{x:.2e}" for x in norms))
    assert all(np.isfinite(norms))
    assert norms[-1] <= 1e-4 * norms[0]
    assert norms[3] < norms[0] and norms[6] < norms[3] and norms[11] < norms[6]


def test_native_autotuner(cuda):
    import torch

    b = _batch()
    u = b.state(seed=4)
    best, table = b.disc.autotune_native(u, reps=2)
    print("autotune", best, table)
    assert table and best in [t[:3] for t in table]
    J1, d1 = b.disc.assemble(u, jprec=F32)
    path = best[0]
    others = [t[:3] for t in table if t[0] == path and t[:3] != best]
    if others:
        b.disc.set_assembly_path(path)
        b.disc.set_launch(others[0][1], others[0][2])
        J2, d2 = b.disc.assemble(u, jprec=F32)
        assert torch.equal(J1, J2) and torch.equal(d1, d2)
    # tetrahedra: the row-lane gather cannot hold their incidences (skipped)
    mesh = M.cavity_3d(4, M.TET, 0.05, seed=2)
    disc = NavierStokesFV1("u,v,w,p", "Inner")
    disc.set_stabilization("fields")
    disc.set_upwind("lps")
    disc.set_kinematic_viscosity(0.01)
    disc.set_mesh(mesh)
    ut = torch.from_numpy(M.random_state(mesh, 3)).to(cuda)
    best, table = disc.autotune_native(ut, reps=1)
    assert all(t[0] != 2 for t in table) and best in [t[:3] for t in table]


def test_misuse_is_refused(cuda):
    import torch

    b = _batch()
    u = b.state(seed=1)
    with pytest.raises(NSFV1Error) as e:
        b.disc.newton_graph_capture(u, g=b.g, krylov=10, stream=None)
    assert e.value.name == "ARGUMENT"
    with pytest.raises(NSFV1Error) as e:
        b.disc.newton_graph_replay()
    assert e.value.name == "ARGUMENT"
    for k in (0, 61):
        with pytest.raises(NSFV1Error) as e:
            b.disc.newton_step(u, g=b.g, krylov=k)
        assert e.value.name == "ARGUMENT"
    with pytest.raises(NSFV1Error) as e:
        b.disc.newton_step(u, krylov=10)  # Dirichlet rows set, values missing
    assert e.value.name == "ARGUMENT"
