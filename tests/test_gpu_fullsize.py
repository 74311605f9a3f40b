"""Full-size workloads, FP32 Jacobian + FP64 defect, time-dependent
stabilisation (u_old, dt = 0.01) and the BDF2 mass part (s_m = 1.5/dt), as
in one Newton iterate of a time step:
- "cavity": the 128^3 hexahedral lid-driven cavity of the benchmark
  (2,097,152 elements, nnz 913,066,000), skewed upwind + FIELDS (split path);
- "channel": a 96x48x48 x 6-tetrahedra obstacle channel like the learned
  predictor's (two obstacle boxes, 1,308,672 elements), LPS upwind + FIELDS
  (split path), at a perturbed parabolic state;
- the batched leg's union of 256 cavities (3,958,533 quadrilaterals, full
  upwind + FIELDS, per-instance viscosity), steady.

- every entry of the 128^3 cavity against the oracle: the oracle assembles
  the mesh in slabs of node planes (each slab with the element layers around
  it, on a thread pool), and every Jacobian entry and defect entry of each
  slab's rows must match to the FP32 / FP64 entry tolerance below.
- sampled rows against the oracle: every element around 300 seeded nodes
  (plus corner, edge and face nodes of the cavity, and inflow, outflow, wall
  and obstacle nodes of the channel) is assembled by the oracle, which
  makes those nodes' rows complete. The GPU defect rows must match to the
  FP64 tolerance, |a-b| <= 1e-12|b| + 1e-13 S; every Jacobian entry of those
  rows to the FP32 entry tolerance |a-b| <= 2^-17 S + 2^-20 |b|; and the GPU
  row products J x (FP32 J, FP64 x) to that tolerance summed over the row:
  sum_k (2^-17 S_k + 2^-20 |J_k|) |x_k|.
- conservation: each SCVF flux enters its two nodes' rows with opposite
  signs, and with no boundary discs nothing else enters the A part. So the
  sum over all nodes of each defect component vanishes up to FP64 rounding,
  and so does each component's row sum of J x for any x (the Jacobian's
  column sums per component vanish) up to FP32 rounding.
- bitwise repeatability of the whole assembly."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from plugin_navierstokes_amd import F32, NavierStokesFV1, PART_A, PART_M
from plugin_navierstokes_amd import meshes as M
from tests._parity import FP32_ABS, FP32_REL, fp32_violation, fp64_violation, worst

pytestmark = pytest.mark.gpu

N = 128
DT = 0.01
S_M = 1.5 / DT
PARTS = PART_A | PART_M


@pytest.fixture(scope="module", params=["cavity", "channel"])
def workload(request, cuda):
    import torch

    if request.param == "cavity":
        et, up, nu = M.HEX, "skewed", 1.0 / 400.0
        mesh = M.cavity_3d(N, M.HEX)
        u = M.cavity_lid_state(mesh, 1.0, base=M.divergence_free_perturbation(mesh, 1e-3, seed=0))
    else:
        et, up, nu = M.TET, "lps", 1.0 / 300.0
        mesh = M.channel_3d((96, 48, 48), n_boxes=2, seed=100)
        u = M.channel_initial_state(mesh, M.channel_dirichlet(mesh)) + 0.05 * M.random_state(mesh, 7)
    uold = 0.999 * u
    disc = NavierStokesFV1("u,v,w,p", "Inner")
    disc.set_stabilization("fields")
    disc.set_upwind(up)
    disc.set_kinematic_viscosity(nu)
    disc.set_mesh(mesh)
    ut, uot = torch.from_numpy(u).to(cuda), torch.from_numpy(uold).to(cuda)
    print(f"full-size {request.param}: ne={mesh.ne} nnz={disc.nnz()}", flush=True)
    J, d = disc.assemble(ut, uot, DT, s_m=S_M, parts=PARTS, jprec=F32)
    params = O.Params(dim=3, upwind=up, nu=nu)
    return request.param, et, params, mesh, disc, u, uold, ut, uot, J, d


def _sample_nodes(name, mesh, count=300, seed=5):
    rng = np.random.default_rng(seed)
    used = np.setdiff1d(np.arange(mesh.nv), mesh.boundary.get("unused", np.zeros(0, np.int64)))
    picks = set(rng.choice(used, size=count, replace=False).tolist())
    if name == "cavity":
        # corners, edge midpoints and face centres of the unit cube
        assert (N + 1) ** 3 == mesh.nv
        for i, j, k in [(0, 0, 0), (N, N, N), (N // 2, 0, 0), (0, N // 2, N), (N // 2, N // 2, 0),
                        (N, N // 2, N // 2), (N // 2, N, N // 2)]:
            for v in np.flatnonzero((np.abs(mesh.xyz - np.array([i, j, k]) / N) < 0.25 / N).all(axis=1)):
                picks.add(int(v))
    else:
        for group in ("inflow", "outflow", "wall", "obstacle"):
            nodes = mesh.boundary[group]
            picks.update(rng.choice(nodes, size=min(8, len(nodes)), replace=False).tolist())
    return np.array(sorted(picks), dtype=np.int64)


def test_fullsize_sampled_rows_vs_oracle(cuda, workload):
    import torch

    name, et, p, mesh, disc, u, uold, _, _, J, d = workload
    nodes = _sample_nodes(name, mesh)
    conn_sub = mesh.conn[np.isin(mesh.conn, nodes).any(axis=1)]
    rp, ci = O.pattern(et, mesh.nv, conn_sub)
    Jr, dr = O.assemble(p, et, mesh.xyz, conn_sub, u, uold, DT, s_m=S_M, parts=PARTS, rowptr=rp, colidx=ci)
    Ja, da = O.assemble_abs(p, et, mesh.xyz, conn_sub, u, uold, DT, s_m=S_M, parts=PARTS, rowptr=rp, colidx=ci)
    rows = (nodes[:, None] * 4 + np.arange(4)[None, :]).ravel()
    rows_t = torch.from_numpy(rows).to(cuda)

    dg = d[rows_t].cpu().numpy()
    assert fp64_violation(dg, dr[rows], da[rows]) <= 1.0, worst(dg, dr[rows], da[rows])

    # every Jacobian entry of the sampled rows: the oracle rows are complete
    # (all incident elements assembled), so their columns are the GPU's
    rpg = disc.rowptr()
    idx = np.concatenate([np.arange(rpg[r], rpg[r + 1]) for r in rows])
    idx_o = np.concatenate([np.arange(rp[r], rp[r + 1]) for r in rows])
    assert len(idx) == len(idx_o)
    Jg = J[torch.from_numpy(idx).to(cuda)].cpu().numpy().astype(np.float64)
    v = fp32_violation(Jg, Jr[idx_o], Ja[idx_o])
    print(f"{name}: {len(idx)} Jacobian entries of {len(rows)} sampled rows, max violation {v:.3g}", flush=True)
    assert v <= 1.0, worst(Jg, Jr[idx_o], Ja[idx_o], FP32_REL, FP32_ABS)

    gen = torch.Generator(cuda).manual_seed(17)
    x = torch.randn(disc.num_dofs(), dtype=torch.float64, device=cuda, generator=gen)
    y = torch.empty_like(x)
    disc.spmv(J, x, y, jprec=F32)
    yg = y[rows_t].cpu().numpy()
    xh = x.cpu().numpy()
    yr = np.empty(len(rows))
    tol = np.empty(len(rows))
    for i, r in enumerate(rows):
        k0, k1 = rp[r], rp[r + 1]
        xs = xh[ci[k0:k1]]
        yr[i] = np.dot(Jr[k0:k1], xs)
        tol[i] = np.dot(FP32_ABS * Ja[k0:k1] + FP32_REL * np.abs(Jr[k0:k1]), np.abs(xs)) + 1e-300
    v = np.abs(yg - yr) / tol
    i = int(np.argmax(v))
    assert v[i] <= 1.0, f"row {rows[i]}: got {yg[i]!r} want {yr[i]!r} tol {tol[i]:.3g}"


def _slab(args):
    """Oracle rows of the node planes [k0, k1) of the cavity: the element
    layers [k0-1, k1) make those rows complete."""
    p, mesh, u, uold, k0, k1 = args
    plane, layer = (N + 1) ** 2, N * N
    conn_sub = mesh.conn[max(k0 - 1, 0) * layer:min(k1, N) * layer]
    rp, ci = O.pattern(O.HEX, mesh.nv, conn_sub)
    Jr, dr = O.assemble(p, O.HEX, mesh.xyz, conn_sub, u, uold, DT, s_m=S_M, parts=PARTS, rowptr=rp, colidx=ci)
    Ja, da = O.assemble_abs(p, O.HEX, mesh.xyz, conn_sub, u, uold, DT, s_m=S_M, parts=PARTS, rowptr=rp, colidx=ci)
    r0, r1 = k0 * plane * 4, k1 * plane * 4
    a, b = rp[r0], rp[r1]
    return k0, k1, ci[a:b].copy(), Jr[a:b].copy(), Ja[a:b].copy(), dr[r0:r1].copy(), da[r0:r1].copy()


def test_fullsize_cavity_every_entry(cuda, workload):
    """All 913,066,000 Jacobian entries and 8,586,756 defect entries of the
    benchmark cavity against the oracle, slab by slab."""
    import os
    from concurrent.futures import ThreadPoolExecutor

    name, _, p, mesh, disc, u, uold, _, _, J, d = workload
    if name != "cavity":
        pytest.skip("every-entry check covers the benchmark cavity (the channel is checked on sampled rows)")
    rpg, cig = disc.pattern()
    dg_all = d.cpu().numpy()
    slabs = [(k, min(k + 8, N + 1)) for k in range(0, N + 1, 8)]
    workers = max(1, min(16, os.cpu_count() or 1))
    worst_j = worst_d = 0.0
    nj = 0
    with ThreadPoolExecutor(max_workers=workers) as ex:
        for k0, k1, ci, Jr, Ja, dr, da in ex.map(_slab, [(p, mesh, u, uold, a, b) for a, b in slabs]):
            r0, r1 = k0 * (N + 1) ** 2 * 4, k1 * (N + 1) ** 2 * 4
            a, b = int(rpg[r0]), int(rpg[r1])
            assert b - a == len(Jr) and np.array_equal(cig[a:b], ci), (k0, k1)
            Jg = J[a:b].cpu().numpy().astype(np.float64)
            vj = fp32_violation(Jg, Jr, Ja)
            vd = fp64_violation(dg_all[r0:r1], dr, da)
            assert vj <= 1.0, (k0, k1, worst(Jg, Jr, Ja, FP32_REL, FP32_ABS))
            assert vd <= 1.0, (k0, k1, worst(dg_all[r0:r1], dr, da))
            worst_j, worst_d = max(worst_j, vj), max(worst_d, vd)
            nj += b - a
            print(f"cavity planes [{k0}, {k1}): {b - a} J entries, violations J {vj:.3g} d {vd:.3g}", flush=True)
    assert nj == disc.nnz() == 913_066_000
    print(f"cavity: every entry checked, worst violations J {worst_j:.3g} d {worst_d:.3g}", flush=True)


def test_fullsize_conservation(cuda, workload):
    """On the A part alone (the mass part is not a flux)."""
    import torch

    disc, ut, uot = workload[4], workload[7], workload[8]
    J, d = disc.assemble(ut, uot, DT, parts=PART_A, jprec=F32)
    dd = d.view(-1, 4)
    s = dd.sum(dim=0).cpu().numpy()
    a = dd.abs().sum(dim=0).cpu().numpy()
    print(f"defect component sums {s} of {a}", flush=True)
    assert np.all(np.abs(s) <= 1e-12 * a)

    gen = torch.Generator(cuda).manual_seed(23)
    x = torch.randn(disc.num_dofs(), dtype=torch.float64, device=cuda, generator=gen)
    y = torch.empty_like(x)
    disc.spmv(J, x, y, jprec=F32)
    z = torch.empty_like(x)
    disc.spmv(J.abs(), x.abs(), z, jprec=F32)
    s = y.view(-1, 4).sum(dim=0).cpu().numpy()
    a = z.view(-1, 4).sum(dim=0).cpu().numpy()
    print(f"row sums of J x {s} of {a}", flush=True)
    assert np.all(np.abs(s) <= 1e-6 * a)


def test_fullsize_repeatable(cuda, workload):
    import torch

    disc, ut, uot, J, d = workload[4], workload[7], workload[8], workload[9], workload[10]
    J2, d2 = disc.assemble(ut, uot, DT, s_m=S_M, parts=PARTS, jprec=F32)
    assert torch.equal(J, J2) and torch.equal(d, d2)


@pytest.fixture(scope="module")
def batch(cuda):
    from plugin_navierstokes_amd.batch import BatchedCavities, draw_instances

    b = BatchedCavities(draw_instances(256, seed=0))
    u = b.state(seed=1)
    print(f"full-size batch: 256 cavities, ne={b.elements} nnz={b.disc.nnz()}", flush=True)
    J, d = b.disc.assemble(u, jprec=F32)
    return b, u, J, d


def test_fullsize_batch_sampled_rows_vs_oracle(cuda, batch):
    """The 256-cavity union (one launch, per-instance viscosity): rows of 48
    seeded nodes, each checked against the oracle run on its own instance's
    elements with that instance's viscosity."""
    import torch

    b, u, J, d = batch
    mesh = b.mesh
    uh = u.cpu().numpy()
    rng = np.random.default_rng(9)
    nodes = np.sort(rng.choice(mesh.nv, size=48, replace=False))
    inst = np.searchsorted(b.noff, nodes, side="right") - 1
    gen = torch.Generator(cuda).manual_seed(31)
    x = torch.randn(b.disc.num_dofs(), dtype=torch.float64, device=cuda, generator=gen)
    y = torch.empty_like(x)
    b.disc.spmv(J, x, y, jprec=F32)
    xh, yh, dh = x.cpu().numpy(), y.cpu().numpy(), d.cpu().numpy()
    rpg = b.disc.rowptr()
    for k in np.unique(inst):
        vs = nodes[inst == k]
        e0, e1 = b.eoff[k], b.eoff[k + 1]
        conn_k = mesh.conn[e0:e1]
        conn_sub = conn_k[np.isin(conn_k, vs).any(axis=1)]
        p = O.Params(dim=2, upwind="full", nu=b.instances[k].nu)
        rp, ci = O.pattern(O.QUAD, mesh.nv, conn_sub)
        Jr, dr = O.assemble(p, O.QUAD, mesh.xyz, conn_sub, uh, parts=PART_A, rowptr=rp, colidx=ci)
        Ja, da = O.assemble_abs(p, O.QUAD, mesh.xyz, conn_sub, uh, parts=PART_A, rowptr=rp, colidx=ci)
        rows = (vs[:, None] * 3 + np.arange(3)[None, :]).ravel()
        assert fp64_violation(dh[rows], dr[rows], da[rows]) <= 1.0, (k, worst(dh[rows], dr[rows], da[rows]))
        # every Jacobian entry of these rows
        idx = np.concatenate([np.arange(rpg[r], rpg[r + 1]) for r in rows])
        idx_o = np.concatenate([np.arange(rp[r], rp[r + 1]) for r in rows])
        assert len(idx) == len(idx_o)
        Jg = J[torch.from_numpy(idx).to(cuda)].cpu().numpy().astype(np.float64)
        assert fp32_violation(Jg, Jr[idx_o], Ja[idx_o]) <= 1.0, (k, worst(Jg, Jr[idx_o], Ja[idx_o], FP32_REL, FP32_ABS))
        for r in rows:
            k0, k1 = rp[r], rp[r + 1]
            xs = xh[ci[k0:k1]]
            want = np.dot(Jr[k0:k1], xs)
            tol = np.dot(FP32_ABS * Ja[k0:k1] + FP32_REL * np.abs(Jr[k0:k1]), np.abs(xs)) + 1e-300
            assert abs(yh[r] - want) <= tol, (k, r, yh[r], want, tol)


def test_fullsize_batch_conservation(cuda, batch):
    import torch

    b, _, J, d = batch
    dd = d.view(-1, 3)
    s = dd.sum(dim=0).cpu().numpy()
    a = dd.abs().sum(dim=0).cpu().numpy()
    print(f"defect component sums {s} of {a}", flush=True)
    assert np.all(np.abs(s) <= 1e-12 * a)
    gen = torch.Generator(cuda).manual_seed(37)
    x = torch.randn(b.disc.num_dofs(), dtype=torch.float64, device=cuda, generator=gen)
    y = torch.empty_like(x)
    b.disc.spmv(J, x, y, jprec=F32)
    z = torch.empty_like(x)
    b.disc.spmv(J.abs(), x.abs(), z, jprec=F32)
    s = y.view(-1, 3).sum(dim=0).cpu().numpy()
    a = z.view(-1, 3).sum(dim=0).cpu().numpy()
    print(f"row sums of J x {s} of {a}", flush=True)
    assert np.all(np.abs(s) <= 1e-6 * a)
