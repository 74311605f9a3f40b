"""Benchmark of the FV1 Navier-Stokes assembly hot path on one MI355X.

Workload: 3D lid-driven cavity on the unit cube, 128^3 Q1 hexahedra
(2,097,152 elements, 2,146,689 nodes), Re = 400 (nu = 2.5e-3), skewed upwind
+ FIELDS stabilisation (RAW diffusion length), time-dependent stabilisation
with the previous time level (dt = 0.01), FP32 Jacobian + FP64 defect.  One
step = one whole-mesh Jacobian + defect assembly of a BDF2 Newton iterate:
stiffness/convection part plus the mass part weighted 1.5/dt, with the state
resident in HBM.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "elem/s Jacobian+defect assembly; s/timestep learned vs plain at 1e-8 residual"
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(nv, ne, nnz, dim, nsh, jbytes, T):
    """Compulsory HBM bytes of one assembly: coordinates, T solution levels,
    connectivity, Jacobian values, defect (CSR indices excluded)."""
    return nv * dim * 8 + T * nv * (dim + 1) * 8 + ne * nsh * 4 + nnz * jbytes + nv * (dim + 1) * 8


DT = 0.01
S_M = 1.5 / DT  # BDF2 weight of the mass part


def host_cores() -> int:
    """Host cores this process may use, capped at 16 (one GPU's share of the
    host on the GPU machines)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(m: int = 48, repeats: int = 5):
    """The CPU oracle (C, -O3 -ffp-contract=off, FP64, scalar CSR scatter) on
    an m^3 sub-cavity of the same workload (same element size class, state
    construction and configuration, A + M parts): 1 thread, as UG4 assembles
    within a process, and every host core with the element-coloured OpenMP
    restatement (standing in for UG4's one MPI rank per core).  Median of
    `repeats` runs after one warm-up run each."""
    import oracle as O
    from plugin_navierstokes_amd import meshes as M

    mesh = M.cavity_3d(m, M.HEX)
    u = M.cavity_lid_state(mesh, 1.0, base=M.divergence_free_perturbation(mesh, 1e-3, seed=0))
    uold = 0.999 * u
    p = O.Params(dim=3, upwind="skewed", nu=1.0 / 400.0)
    rp, ci = O.pattern(O.HEX, mesh.nv, mesh.conn)
    order, cptr = O.structured_colors(mesh.shape)
    cores = host_cores()

    def timed(fn):
        fn()
        ts = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    t1 = timed(lambda: O.assemble(p, O.HEX, mesh.xyz, mesh.conn, u, uold, DT, s_m=S_M, parts=3, rowptr=rp,
                                  colidx=ci))
    tn = timed(lambda: O.assemble_colored(p, O.HEX, mesh.xyz, mesh.conn, u, order, cptr, uold, DT, s_m=S_M, parts=3,
                                          rowptr=rp, colidx=ci, threads=cores))
    sample = (f"CPU oracle (C, -O3 -ffp-contract=off, FP64), Jacobian+defect assembly (A + M parts) of a {m}^3-hexahedron "
              f"cavity ({mesh.ne} elements, same element size class, state and configuration); median of {repeats} "
              f"runs after a warm-up")
    return {"value": mesh.ne / t1, "unit": "elem/s", "cores": 1, "kind": "port", "sample": sample,
            "seconds_per_run": t1,
            "all_cores": {"value": mesh.ne / tn, "unit": "elem/s", "cores": cores, "seconds_per_run": tn,
                          "method": "element colouring (8 colours), OpenMP over the elements of a colour"}}


def parity_sample(mesh, u, uold, J, d, rowptr_full, nodes=64, seed=5):
    """Every Jacobian and defect entry of the rows of `nodes` seeded nodes
    against the FP64 oracle (all elements around them assembled): the
    largest per-entry error relative to the entry's scale S (sum of the
    magnitudes of its element contributions) and the largest ratio to the
    stated bound (<= 1 passes)."""
    import torch

    import oracle as O

    rng = np.random.default_rng(seed)
    vs = np.sort(rng.choice(mesh.nv, size=nodes, replace=False))
    conn_sub = mesh.conn[np.isin(mesh.conn, vs).any(axis=1)]
    p = O.Params(dim=3, upwind="skewed", nu=1.0 / 400.0)
    rp, ci = O.pattern(O.HEX, mesh.nv, conn_sub)
    Jr, dr = O.assemble(p, O.HEX, mesh.xyz, conn_sub, u, uold, DT, s_m=S_M, parts=3, rowptr=rp, colidx=ci)
    Ja, da = O.assemble_abs(p, O.HEX, mesh.xyz, conn_sub, u, uold, DT, s_m=S_M, parts=3, rowptr=rp, colidx=ci)
    rows = (vs[:, None] * 4 + np.arange(4)[None, :]).ravel()
    idx = np.concatenate([np.arange(rowptr_full[r], rowptr_full[r + 1]) for r in rows])
    idx_o = np.concatenate([np.arange(rp[r], rp[r + 1]) for r in rows])
    Jg = J[torch.from_numpy(idx).to(J.device)].cpu().numpy().astype(np.float64)
    dg = d[torch.from_numpy(rows).to(d.device)].cpu().numpy()
    Jo, So, do, Sd = Jr[idx_o], Ja[idx_o], dr[rows], da[rows]
    ej = np.abs(Jg - Jo)
    ed = np.abs(dg - do)
    return {"rows": int(len(rows)), "entries_J": int(len(idx)),
            "max_err_J": float(np.max(ej / np.maximum(So, 1e-300))),
            "max_err_d": float(np.max(ed / np.maximum(Sd, 1e-300))),
            "bound": "J (FP32): |a-b| <= 2^-17 S + 2^-20 |b|; d (FP64): |a-b| <= 1e-12 |b| + 1e-13 S",
            "max_violation_J": float(np.max(ej / (2.0 ** -17 * So + 2.0 ** -20 * np.abs(Jo) + 1e-300))),
            "max_violation_d": float(np.max(ed / (1e-12 * np.abs(do) + 1e-13 * Sd + 1e-300))),
            "reference": "FP64 CPU oracle, same state"}


def timestep_leg(mesh, disc, steps: int, dt: float = 0.01):
    """Plain GPU time stepping of the same cavity: BDF1 start (untimed) then
    `steps` timed BDF2 steps, each solved by Newton (FP32 Jacobian as the
    preconditioner, FP64 defect and update, GMRES + geometric multigrid) to
    a 1e-8 relative FP64 residual (relative to the step system's defect at
    the old state), initial guess = linear extrapolation."""
    import torch

    from plugin_navierstokes_amd import F32
    from plugin_navierstokes_amd import meshes as M
    from plugin_navierstokes_amd.solver import BDFStepper, NewtonSolver, build_multigrid, cavity_dirichlet

    dofs, vals = cavity_dirichlet(mesh)
    # one undamped block-Jacobi sweep before and after the coarse correction:
    # as a GMRES preconditioner this is cheaper per iteration than 2 + 2
    # damped sweeps and takes few more iterations
    levels = build_multigrid(disc, mesh, cavity_dirichlet, pre=1, post=1, omega=1.0)
    # the Newton iteration contracts linearly (about 10x per iteration: the
    # FV1 Jacobian omits the derivatives of the upwind and stabilisation
    # weights), so correction solves to 3e-2 take as many Newton iterations
    # as tighter ones and fewer GMRES iterations in total
    nt = NewtonSolver(disc, dofs, vals, jprec=F32, rtol=1e-8, atol=0.0, maxit=30, lin_rtol=3e-2, lin_maxit=200,
                      restart=30, multigrid=True, freeze_upwind_after=3)
    stepper = BDFStepper(nt, dt, reference="old", mg_lag=True)
    u = torch.from_numpy(M.cavity_lid_state(mesh, 1.0, M.divergence_free_perturbation(mesh, 1e-3, seed=0))).cuda()
    u, st0 = stepper.step(u)  # BDF1 start, untimed
    torch.cuda.synchronize()
    times, its, lin, fresh = [], [], [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        u, st = stepper.step(u, guess=stepper.extrapolate(u))
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        its.append(st.iterations)
        lin.append(st.linear_iterations)
        fresh.append((st.fresh_residual if st.fresh_residual is not None else st.residuals[-1]) / st.reference)
        if not st.converged:
            raise RuntimeError(f"time step did not converge: residuals {st.residuals}")
        print(f"bench: BDF2 step {len(times)}: {times[-1]:.3f} s, newton {st.iterations}, gmres {st.linear_iterations}",
              file=sys.stderr, flush=True)
    return {"method": "plain (linear extrapolation guess)", "s_per_step": float(np.mean(times)),
            "newton_per_step": float(np.mean(its)), "gmres_per_step": float(np.mean(lin)), "steps": steps,
            "levels": len(levels) + 1, "residual": "1e-8 x ||d(u^n)||, FP64", "correction_rtol": nt.lin_rtol,
            "smoothing": "1 + 1 block-Jacobi sweeps, omega 1.0",
            "upwind_decisions": "skewed corners decided anew for 2 Newton iterations, then replayed",
            "residual_kind": ("1e-8 is reached on the frozen-decision system (the skewed corner choices of the third "
                              "iterate replayed); with the corners decided anew at the converged state the defect "
                              "is max_fresh_residual_ratio of the reference norm"),
            "max_fresh_residual_ratio": float(np.max(fresh)),
            "workload": f"{mesh.shape[0]}^3 hex cavity, Re=400, skewed+FIELDS, BDF2 dt={dt}, FP32 J + FP64 defect"}


def learned_leg(steps: int):
    """Plain (linear extrapolation) vs learned initial guess on a held-out
    96x48x48x6-tetrahedra channel (Re, obstacles and seed not in the
    training set), with the quadratic extrapolation alone (no network) as
    the ablation: BDF2 dt=0.005, LPS + FIELDS, FP32 J + FP64 defect, 1e-8
    relative FP64 residual; 6 untimed warm-up steps, then `steps` timed
    steps per method.  The channel assembly's roofline uses its autotuned
    launch time."""
    from plugin_navierstokes_amd.learned import compare

    w = os.path.join(ROOT, "plugin_navierstokes_amd", "data", "predictor_channel.npz")
    r = compare(w, seeds=(100,), steps=steps, warm=6, log=lambda m: print("bench: " + m, file=sys.stderr, flush=True))
    r["workload"] = ("channel 96x48x48 hexahedra x 6 tetrahedra (1.26M elements after obstacles), held-out seed 100, "
                     "BDF2 dt=0.005, LPS+FIELDS, FP32 J + FP64 defect, residual 1e-8 x ||d(u^n)||")
    r["learned_vs_plain"] = r["learned_s_per_step"] / r["plain_s_per_step"]
    r["quadratic_vs_plain"] = r["quadratic_s_per_step"] / r["plain_s_per_step"]
    launch = r.get("launch") or {}
    geo = r.pop("mesh_sizes", None)
    if geo and launch.get("ms"):
        B = algorithmic_bytes(geo["nv"], geo["ne"], geo["nnz"], 3, 4, 4, 2)
        a = B / (launch["ms"] * 1e-3) / 1e9
        r["roofline"] = {"bound": "hbm", "achieved": a, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": a / HBM_PEAK_GBS,
                         "algorithmic_bytes": B, "kernel_ms": launch["ms"],
                         "kernel": "one channel assembly (A + M parts, FP32 J), autotuned launch"}
    return r


def channel_assembly_leg(steps: int, launch=None):
    """The assembly alone on the held-out channel of the learned-guess
    comparison (96x48x48 hexahedra x 6 tetrahedra minus obstacles, seed 100,
    its Re and obstacle count drawn as there): LPS + FIELDS, time-dependent
    stabilisation (dt = 0.005), A + M parts, FP32 J + FP64 defect, on the
    initial channel state.  Launch shape from the autotuner unless given;
    `steps` timed assemblies (HIP events on the launch stream)."""
    import torch

    from plugin_navierstokes_amd import F32, NavierStokesFV1, PART_A, PART_M
    from plugin_navierstokes_amd import meshes as M
    from plugin_navierstokes_amd.batch import autotune

    seed, dt = 100, 0.005
    rng = np.random.default_rng(1000 + seed)
    re = float(rng.uniform(100.0, 500.0))
    nbox = int(rng.integers(1, 5))
    mesh = M.channel_3d((96, 48, 48), n_boxes=nbox, seed=seed)
    disc = NavierStokesFV1("u,v,w,p", "Inner")
    disc.set_stabilization("fields")
    disc.set_upwind("lps")
    disc.set_kinematic_viscosity(1.0 / re)
    disc.set_mesh(mesh)
    dofs, vals = M.channel_dirichlet(mesh)
    u = torch.from_numpy(M.channel_initial_state(mesh, (dofs, vals))).to("cuda:0")
    J = torch.empty(disc.nnz(), dtype=torch.float32, device="cuda:0")
    d = torch.empty(disc.num_dofs(), dtype=torch.float64, device="cuda:0")
    stream = torch.cuda.current_stream()

    def step():
        disc.assemble(u, u, dt, s_m=1.5 / dt, parts=PART_A | PART_M, jprec=F32, J=J, d=d, stream=stream, check=False)

    if launch:
        disc.set_assembly_path(launch[0])
        disc.set_launch(launch[1], launch[2])
        shape = {"path": launch[0], "threads": launch[1], "lds_bytes": launch[2], "autotuned": False}
    else:
        best, _ = autotune(disc, step, reps=2)
        shape = {"path": best[0], "threads": best[1], "lds_bytes": best[2], "autotuned": True}
    for _ in range(3):
        step()
    disc.sync_errors(stream)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for a, b in ev:
        a.record(stream)
        step()
        b.record(stream)
    torch.cuda.synchronize()
    disc.sync_errors(stream)
    ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    B = algorithmic_bytes(mesh.nv, mesh.ne, disc.nnz(), 3, 4, 4, 2)
    a = B / (ms * 1e-3) / 1e9
    return {"workload": f"held-out channel (seed {seed}, Re {re:.1f}, {nbox} obstacles): {mesh.ne} tetrahedra, "
                        f"{mesh.nv} nodes, nnz {disc.nnz()}, LPS + FIELDS, dt {dt}, A + M parts, FP32 J + FP64 defect",
            "ms_per_assembly": ms, "elem_per_s": mesh.ne / (ms * 1e-3), "path_used": disc.assembly_path_used(),
            "launch": shape,
            "roofline": {"bound": "hbm", "achieved": a, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": a / HBM_PEAK_GBS,
                         "algorithmic_bytes": B, "kernel_ms": ms,
                         "kernel": CHANNEL_KERNELS.get(disc.assembly_path_used(), "one assembly")}}


def coupled_leg(steps: int, n: int = 64):
    """The coupled path (csrc/nsfv1_coupled.hip: SCVFs of an element solved
    together) on an n^3 hexahedral cavity with the 128^3 leg's state
    construction, Re = 400, FP32 J + FP64 defect, A part: the FLOW
    stabilisation with the LPS upwind, and FIELDS with the positive upwind.
    `steps` timed assemblies each (HIP events), after one untimed one."""
    import torch

    from plugin_navierstokes_amd import F32, NavierStokesFV1, PART_A
    from plugin_navierstokes_amd import meshes as M

    mesh = M.cavity_3d(n, M.HEX)
    u = torch.from_numpy(M.cavity_lid_state(mesh, 1.0, base=M.divergence_free_perturbation(mesh, 1e-3, seed=0))).to(
        "cuda:0")
    out = {}
    for name, stab, upw in (("flow_lps", "flow", "lps"), ("fields_positive", "fields", "positive")):
        disc = NavierStokesFV1("u,v,w,p", "Inner")
        disc.set_stabilization(stab)
        disc.set_upwind(upw)
        disc.set_kinematic_viscosity(1.0 / 400.0)
        disc.set_mesh(mesh)
        J = torch.empty(disc.nnz(), dtype=torch.float32, device="cuda:0")
        d = torch.empty(disc.num_dofs(), dtype=torch.float64, device="cuda:0")
        stream = torch.cuda.current_stream()
        disc.assemble(u, parts=PART_A, jprec=F32, J=J, d=d, stream=stream)
        disc.sync_errors(stream)
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        for a, b in ev:
            a.record(stream)
            disc.assemble(u, parts=PART_A, jprec=F32, J=J, d=d, stream=stream, check=False)
            b.record(stream)
        torch.cuda.synchronize()
        disc.sync_errors(stream)
        ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
        B = algorithmic_bytes(mesh.nv, mesh.ne, disc.nnz(), 3, 8, 4, 1)
        ach = B / (ms * 1e-3) / 1e9
        out[name] = {"ms_per_assembly": ms, "elem_per_s": mesh.ne / (ms * 1e-3), "path_used": disc.assembly_path_used(),
                     "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": ach / HBM_PEAK_GBS, "algorithmic_bytes": B, "kernel_ms": ms,
                                  "kernel": "one assembly = nsfv1::coupled_scvf<3,float> + nsfv1::coupled_rows<3,float>"}}
        del disc, J, d
        torch.cuda.empty_cache()
    out["workload"] = (f"{n}^3 hex cavity ({mesh.ne} elements), Re = 400, FP32 J + FP64 defect, A part; FLOW + LPS "
                       "and FIELDS + positive upwind (coupled path)")
    return out


def batched_leg(steps: int = 20, count: int = 256):
    """256 cavity instances (side ~ U{32..192}, Re log-uniform [10, 2000],
    lid ~ U[0.5, 1.5], one seed) assembled in one launch: launch shape from
    the C++ autotuner, assembly + Dirichlet rows captured in a HIP graph and
    replayed, and a whole Newton step in the library's own graph."""
    import torch

    from plugin_navierstokes_amd.batch import BatchedCavities, draw_instances

    t0 = time.perf_counter()
    b = BatchedCavities(draw_instances(count, seed=0))
    u = b.state(seed=1)
    print(f"bench: batch of {count} cavities, {b.elements} quads, setup {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    best, table = b.disc.autotune_native(u, reps=3)
    table = [(p, t, l, m) for p, t, l, m in table]
    b.capture(u)
    for _ in range(3):
        b.replay()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(steps):
        b.replay()
    e.record()
    torch.cuda.synchronize()
    ms = s.elapsed_time(e) / steps
    # a whole Newton step in the library's own HIP graph (C ABI): assembly,
    # Dirichlet rows, 20 block-Jacobi GMRES iterations with device-side
    # reductions, update
    un = u.clone()
    r2 = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    b.capture_newton(un, krylov=20, res2=r2)
    for _ in range(2):
        b.replay_newton()
    torch.cuda.synchronize()
    s2, e2 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s2.record(b.nt_stream)
    for _ in range(steps):
        b.replay_newton()
    e2.record(b.nt_stream)
    torch.cuda.synchronize()
    b.disc.sync_errors()
    nms = s2.elapsed_time(e2) / steps
    mesh = b.mesh
    B = algorithmic_bytes(mesh.nv, mesh.ne, b.disc.nnz(), 2, 4, 4, 1)
    a = B / (ms * 1e-3) / 1e9
    return {"workload": f"{count} 2D cavities in one launch (side U{{32..192}}, Re log-U[10,2000], lid U[0.5,1.5], "
                        f"seed 0), full upwind + FIELDS, FP32 J + FP64 defect, assembly + Dirichlet rows per HIP-graph "
                        f"replay",
            "elements": int(b.elements), "ms_per_step": ms, "elem_per_s": b.elements / (ms * 1e-3),
            "launch": {"path": best[0], "threads": best[1], "lds_bytes": best[2]},
            "roofline": {"bound": "hbm", "achieved": a, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": a / HBM_PEAK_GBS,
                         "algorithmic_bytes": B, "kernel_ms": ms,
                         "kernel": "one graph replay: union assembly (steady, FP32 J) + Dirichlet rows"},
            "autotune": [{"path": p, "threads": t, "lds": l, "ms": m} for p, t, l, m in table],
            "newton_step": {"ms_per_step": nms, "krylov": 20, "preconditioner": "point-block Jacobi",
                            "graph": "nsfv1_newton_graph_capture / nsfv1_newton_graph_replay (C ABI)",
                            "contents": "assembly, Dirichlet rows, 20 GMRES iterations with device-side reductions "
                                        "and least squares, FP64 update; no host synchronisation"}}


# the kernels of one 128^3 hex assembly per assembly path that ran
ASSEMBLY_KERNELS = {
    3: "nsfv1::march::march_hex<2, false> (plane march: one launch per assembly)",
    0: "one assembly = nsfv1::scvf_records<3,2,Q> (Q=0,1,2) + nsfv1::assemble_rows_split<3,2>",
    2: "one assembly = nsfv1::scvf_records<3,2,Q> (Q=0,1,2) + nsfv1::assemble_rows_rowlane<3,2>",
    1: "nsfv1::assemble_rows<3,float,2> (fused, one launch per assembly)",
}

# the tetrahedral channel (LPS upwind: template argument 3)
CHANNEL_KERNELS = {
    0: "one assembly = nsfv1::scvf_records<2,3,Q> + nsfv1::assemble_rows_split<2,3> (split path)",
    2: "one assembly = nsfv1::scvf_records<2,3,Q> + nsfv1::assemble_rows_rowlane<2,3>",
    1: "nsfv1::assemble_rows<2,float,3> (fused, one launch per assembly)",
}


def measured_traffic(n: int):
    """HBM bytes per assembly launch from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE passes scaled by a plain-copy calibration);
    None when no summary covers this workload."""
    if n != 128:
        return None
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "assembly_hex128_fp32.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_dispatch"):
                return int(line.split("=")[1])
    except OSError:
        return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--timestep-steps", type=int, default=50,
                    help="timed BDF2 steps of the 128^3 cavity time stepping (0: skip)")
    ap.add_argument("--channel-steps", type=int, default=200,
                    help="timed BDF2 steps per method (plain, learned) of the channel (0: skip)")
    ap.add_argument("--no-batched", dest="batched", action="store_false", help="skip the batched-instances leg")
    ap.add_argument("--no-channel", dest="channel", action="store_false",
                    help="skip the learned-vs-plain channel leg of the time stepping (profiling runs)")
    ap.add_argument("--launch", type=int, nargs=3, metavar=("PATH", "THREADS", "LDS"),
                    help="fixed assembly path / launch shape instead of the autotuner (profiling runs)")
    ap.add_argument("--channel-assembly", type=int, default=20,
                    help="timed assemblies of the held-out channel alone (0: skip)")
    ap.add_argument("--coupled", type=int, default=5,
                    help="timed assemblies of the coupled path (FLOW, positive upwind) on a 64^3 cavity (0: skip)")
    ap.add_argument("--channel-launch", type=int, nargs=3, metavar=("PATH", "THREADS", "LDS"),
                    help="fixed launch shape of the channel assembly leg (profiling runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 1 GiB device copy (rocprof PMC byte calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: measures one GPU only; sharding the assembly over several GPUs is not "
              "implemented, so --gpus must be 1.", file=sys.stderr)
        sys.exit(2)

    import torch

    from plugin_navierstokes_amd import F32, NavierStokesFV1, PART_A, PART_M
    from plugin_navierstokes_amd import meshes as M

    dev = torch.device("cuda:0")
    n = args.n
    t0 = time.perf_counter()
    mesh = M.cavity_3d(n, M.HEX)
    u = M.cavity_lid_state(mesh, 1.0, base=M.divergence_free_perturbation(mesh, 1e-3, seed=0))
    uold = 0.999 * u
    disc = NavierStokesFV1("u,v,w,p", "Inner")
    disc.set_stabilization("fields")
    disc.set_upwind("skewed")
    disc.set_kinematic_viscosity(1.0 / 400.0)
    disc.set_mesh(mesh)
    print(f"bench: mesh {n}^3 hex, nv={mesh.nv} ne={mesh.ne} nnz={disc.nnz()} setup {time.perf_counter()-t0:.1f}s",
          file=sys.stderr, flush=True)
    ut = torch.from_numpy(u).to(dev)
    uot = torch.from_numpy(uold).to(dev)
    J = torch.empty(disc.nnz(), dtype=torch.float32, device=dev)
    d = torch.empty(disc.num_dofs(), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()

    if args.copy_calibration:
        a = torch.empty(1 << 28, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()

    def step():
        disc.assemble(ut, uot, DT, s_m=S_M, parts=PART_A | PART_M, jprec=F32, J=J, d=d, stream=stream, check=False)

    # launch shape (incidences per workgroup, LDS budget) from the autotuner
    from plugin_navierstokes_amd.batch import autotune

    if args.launch:
        disc.set_assembly_path(args.launch[0])
        disc.set_launch(args.launch[1], args.launch[2])
        launch = {"path": args.launch[0], "threads": args.launch[1], "lds_bytes": args.launch[2], "autotuned": False}
    else:
        best, table = autotune(disc, step, reps=2)
        launch = {"path": best[0], "threads": best[1], "lds_bytes": best[2], "autotuned": True,
                  "candidates_ms": [[p, t, l, round(m, 3)] for p, t, l, m in table]}
    for _ in range(args.warmup):
        step()
    disc.sync_errors(stream)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s, e in ev:
        s.record(stream)
        step()
        e.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    disc.sync_errors(stream)
    kms = [s.elapsed_time(e) for s, e in ev]
    ms_step = wall * 1000.0 / args.steps
    kavg = float(np.mean(kms))
    nnz = disc.nnz()
    B = algorithmic_bytes(mesh.nv, mesh.ne, nnz, 3, 8, 4, 2)
    achieved = B / (kavg * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": mesh.ne / (ms_step * 1e-3),
        "unit": "elem/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (Jacobian stored f32)",
        "data": "synthetic",
        "config": {
            "workload": f"3D lid-driven cavity, {n}^3 Q1 hexahedra, Re=400, skewed upwind + FIELDS (RAW), "
                        f"time-dependent stabilisation (u_old, dt=0.01), BDF2 Newton iterate (A part + mass part "
                        f"x 1.5/dt), FP32 Jacobian + FP64 defect",
            "elements": mesh.ne, "nodes": mesh.nv, "nnz": nnz, "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(n),
            "kernel": ASSEMBLY_KERNELS.get(disc.assembly_path_used(), "one assembly"),
            "kernel_ms": kavg, "algorithmic_bytes": B, "launch": launch,
        },
    }
    out["parity"] = parity_sample(mesh, u, uold, J, d, disc.rowptr())
    print(f"bench: parity {out['parity']}", file=sys.stderr, flush=True)
    del J, d
    torch.cuda.empty_cache()
    out["timestep"] = {}
    if args.timestep_steps > 0:
        out["timestep"]["plain"] = timestep_leg(mesh, disc, args.timestep_steps)
    del disc
    torch.cuda.empty_cache()
    if args.channel and args.channel_steps > 0:
        out["timestep"]["channel"] = learned_leg(args.channel_steps)
    if args.channel_assembly > 0:
        out["channel_assembly"] = channel_assembly_leg(args.channel_assembly, args.channel_launch)
        print(f"bench: channel assembly {out['channel_assembly']['ms_per_assembly']:.3f} ms", file=sys.stderr,
              flush=True)
    if args.coupled > 0:
        out["coupled"] = coupled_leg(args.coupled)
        print(f"bench: coupled path {out['coupled']['flow_lps']['ms_per_assembly']:.2f} ms (FLOW + LPS), "
              f"{out['coupled']['fields_positive']['ms_per_assembly']:.2f} ms (FIELDS + positive)", file=sys.stderr,
              flush=True)
    if args.batched:
        out["batched"] = batched_leg()
    if not args.no_cpu:
        # a bounded sample: the 64^3 sub-cavity (262k elements) of the same workload
        out["cpu_baseline"] = cpu_baseline(min(n, 64))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
