"""Batched instances: the union assembly equals the per-instance assemblies
bitwise, the HIP-graph replay equals the eager step on every replay, and
the autotuner's launch shape changes nothing in the results."""
from __future__ import annotations

import numpy as np
import pytest

from plugin_navierstokes_amd import F32, F64, NavierStokesFV1, PART_A
from plugin_navierstokes_amd import meshes as M
from plugin_navierstokes_amd.batch import BatchedCavities, Instance, autotune, draw_instances

pytestmark = pytest.mark.gpu


def _instances():
    return [Instance(n, re, lid, lid / re) for n, re, lid in
            [(4, 10.0, 1.0), (7, 150.0, 0.6), (5, 2000.0, 1.4), (8, 35.0, 0.9), (6, 700.0, 1.2)]]


@pytest.mark.parametrize("jprec", [F64, F32])
def test_union_equals_separate(cuda, jprec):
    import torch

    ins = _instances()
    b = BatchedCavities(ins, jprec=jprec)
    u = b.state(seed=3, scale=0.3)
    b.disc.assemble(u, None, 0.0, parts=PART_A, jprec=jprec, J=b.J, d=b.d)
    rp_u, _ = b.disc.pattern()
    Ju, du = b.J.cpu().numpy(), b.d.cpu().numpy()
    for i, inst in enumerate(ins):
        m = M.cavity_2d(inst.n, M.QUAD)
        disc = NavierStokesFV1("u,v,p", "Inner")
        disc.set_stabilization("fields")
        disc.set_upwind("full")
        disc.set_kinematic_viscosity(inst.nu)
        disc.set_mesh(m)
        ui = u[b.noff[i]:b.noff[i + 1]].contiguous()
        J, d = disc.assemble(ui, jprec=jprec)
        r0, r1 = 3 * b.noff[i], 3 * b.noff[i + 1]
        assert np.array_equal(Ju[rp_u[r0]:rp_u[r1]], J.cpu().numpy()), i
        assert np.array_equal(du[r0:r1], d.cpu().numpy()), i


def test_graph_replay_matches_eager(cuda):
    import torch

    b = BatchedCavities(draw_instances(12, seed=5, side=(8, 24)))
    u = b.state(seed=2)
    b.capture(u)
    for k in range(3):
        with torch.no_grad():
            u.add_(0.01 * torch.sin(torch.arange(u.numel(), device=u.device, dtype=u.dtype).view_as(u) * (k + 1)))
            u.view(-1)[torch.as_tensor(b.dofs, device=u.device)] = b.g
        torch.cuda.synchronize()
        b.replay()
        torch.cuda.synchronize()
        Jg, dg = b.J.clone(), b.d.clone()
        b.step(u)
        torch.cuda.synchronize()
        assert torch.equal(Jg, b.J) and torch.equal(dg, b.d), f"replay {k}"


def test_autotune_keeps_results(cuda):
    """Launch shapes never change results; the two assembly paths agree to
    FP32 rounding (J) and FP64 rounding (d)."""
    import torch

    b = BatchedCavities(draw_instances(10, seed=7, side=(8, 20)))
    u = b.state(seed=4)
    b.step(u)
    J0, d0 = b.J.clone(), b.d.clone()
    best, table = autotune(b.disc, lambda: b.step(u), candidates=[(0, 64, 16384), (1, 128, 32768), (0, 256, 65536),
                                                                  (1, 64, 16384)], reps=2)
    print("autotune", best, table)
    assert len(table) == 4 and best in table
    b.step(u)
    torch.cuda.synchronize()
    J1, d1 = b.J.clone(), b.d.clone()
    b.disc.set_launch(256, 65536)  # same path, another shape: bitwise equal
    b.step(u)
    torch.cuda.synchronize()
    assert torch.equal(J1, b.J) and torch.equal(d1, b.d)
    scale = J0.abs().max()
    assert float((J1 - J0).abs().max()) <= 1e-6 * float(scale)
    assert float((d1 - d0).abs().max()) <= 1e-12 * float(d0.abs().max())


def test_autotune_skips_shapes_a_mesh_cannot_use(cuda):
    """A tetrahedral node has more incidences than the 16 that one 64-lane
    workgroup of the row-lane gather holds: the autotuner skips that shape
    (the library refuses it with NSFV1_ERR_ARGUMENT) and keeps a usable one."""
    import torch

    mesh = M.cavity_3d(4, M.TET, 0.05, seed=2)
    disc = NavierStokesFV1("u,v,w,p", "Inner")
    disc.set_stabilization("fields")
    disc.set_upwind("lps")
    disc.set_kinematic_viscosity(0.01)
    disc.set_mesh(mesh)
    u = torch.from_numpy(M.random_state(mesh, 3)).to(cuda)
    run = lambda: disc.assemble(u, jprec=F32)  # noqa: E731
    best, table = autotune(disc, run, candidates=[(2, 64, 16384), (0, 64, 16384), (0, 128, 32768)], reps=1)
    assert [r[:3] for r in table] == [(0, 64, 16384), (0, 128, 32768)]
    assert best in table
    J, d = run()
    assert torch.isfinite(J).all() and torch.isfinite(d).all()
