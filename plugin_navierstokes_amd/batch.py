"""Batched cavity instances in one assembly launch, its HIP-graph capture
and the launch-shape autotuner.

Many independent 2D cavities of different sizes, Reynolds numbers and lid
speeds are concatenated into one disjoint mesh.  The assembly kernel reads
each element's viscosity from the instance table
(``nsfv1_batch_set_instances``), so one launch assembles every instance;
Dirichlet rows of every instance go through the union's Dirichlet list.
The assembly + Dirichlet step is captured once into a HIP graph (through
torch's graph API on a side stream) and replayed; a whole Newton step
(assembly, Dirichlet rows, device-side GMRES, update) is captured by the
library itself (``nsfv1_newton_graph_capture``), so a C++ caller gets the
same graph.  ``autotune`` times the launch shapes (incidences per
workgroup, LDS budget) on the actual batch and keeps the fastest
(``nsfv1_autotune`` is the same search in C++).  No counterpart in the
reference plugin.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from . import F32, PART_A, NavierStokesFV1
from . import meshes as M
from .solver import cavity_dirichlet


@dataclass
class Instance:
    n: int          # cells per side
    re: float       # Reynolds number (lid speed 1 over the unit length / nu ... nu = lid / re)
    lid: float      # lid speed
    nu: float


def draw_instances(count: int = 256, seed: int = 0, side=(32, 192), re=(10.0, 2000.0), lid=(0.5, 1.5)):
    """Side n ~ U{32..192} (so E[sum of elements] = count * E[n^2]), Re
    log-uniform, lid speed uniform; nu = lid * 1 / Re."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.integers(side[0], side[1] + 1))
        r = float(np.exp(rng.uniform(np.log(re[0]), np.log(re[1]))))
        u = float(rng.uniform(lid[0], lid[1]))
        out.append(Instance(n, r, u, u / r))
    return out


def union_mesh(instances, et=M.QUAD):
    """Disjoint union of unit-square cavities.  Returns (mesh, node_offset,
    elem_offset, dirichlet dofs, dirichlet values)."""
    xyz, conn, dofs, vals = [], [], [], []
    noff, eoff = [0], [0]
    for ins in instances:
        m = M.cavity_2d(ins.n, et)
        d, v = cavity_dirichlet(m, lid_speed=ins.lid)
        xyz.append(m.xyz)
        conn.append(m.conn + noff[-1])
        dofs.append(d + 3 * noff[-1])
        vals.append(v)
        noff.append(noff[-1] + m.nv)
        eoff.append(eoff[-1] + m.ne)
    mesh = M.Mesh(et, np.concatenate(xyz), np.concatenate(conn).astype(np.int32), {}, ())
    dofs = np.concatenate(dofs)
    vals = np.concatenate(vals)
    o = np.argsort(dofs)
    return mesh, np.asarray(noff), np.asarray(eoff), dofs[o].astype(np.int64), vals[o]


class BatchedCavities:
    """One discretisation over the union of all instances (full upwind +
    FIELDS, FP32 Jacobian + FP64 defect by default)."""

    def __init__(self, instances, et=M.QUAD, upwind="full", jprec=F32):
        import torch

        self.instances = instances
        self.mesh, self.noff, self.eoff, self.dofs, self.vals = union_mesh(instances, et)
        self.disc = NavierStokesFV1("u,v,p", "Inner")
        self.disc.set_stabilization("fields")
        self.disc.set_upwind(upwind)
        self.disc.set_mesh(self.mesh)
        self.disc.batch_set_instances(self.eoff, [i.nu for i in instances])
        self.disc.set_dirichlet(self.dofs)
        self.jprec = jprec
        dev = torch.device("cuda:0")
        self.g = torch.as_tensor(self.vals, device=dev)
        self.J = torch.empty(self.disc.nnz(), dtype=torch.float32 if jprec == F32 else torch.float64, device=dev)
        self.d = torch.empty(self.disc.num_dofs(), dtype=torch.float64, device=dev)
        self.graph = None

    @property
    def elements(self) -> int:
        return self.mesh.ne

    def state(self, seed: int = 1, scale: float = 0.1):
        """A seeded state with every instance's lid and walls imposed."""
        import torch

        u = scale * M.random_state(self.mesh, seed)
        u.reshape(-1)[self.dofs] = self.vals
        return torch.from_numpy(u).to("cuda:0")

    def step(self, u, stream=None):
        """Newton-step assembly: J and d of every instance, Dirichlet rows applied."""
        self.disc.assemble(u, None, 0.0, parts=PART_A, jprec=self.jprec, J=self.J, d=self.d, stream=stream,
                           check=False)
        self.disc.apply_dirichlet(self.J, self.jprec, self.d, u, self.g, stream=stream)

    def capture(self, u):
        """Record step(u) into a HIP graph (after an eager warm-up that builds
        the plan and the geometry cache, so nothing is allocated while
        capturing).  u must stay at the same address for replays."""
        import torch

        self.step(u)
        torch.cuda.synchronize()
        s = torch.cuda.Stream()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            self.step(u, stream=torch.cuda.current_stream())
        torch.cuda.synchronize()
        self.graph = g
        return g

    def replay(self):
        self.graph.replay()

    def newton_step(self, u, krylov: int = 20, res2=None, stream=None):
        """One Newton step of every instance on the device
        (nsfv1_newton_step): assembly, Dirichlet rows, `krylov` block-Jacobi
        GMRES iterations with device-side reductions, u -= correction."""
        self.disc.newton_step(u, g=self.g, J=self.J, d=self.d, krylov=krylov, res2=res2, jprec=self.jprec,
                              stream=stream)

    def capture_newton(self, u, krylov: int = 20, res2=None, stream=None):
        """Capture the Newton step into the library's HIP graph
        (nsfv1_newton_graph_capture, through the C ABI); replay with
        replay_newton.  u, res2 must stay at their addresses."""
        import torch

        self.nt_stream = stream or torch.cuda.Stream()
        self.disc.newton_graph_capture(u, g=self.g, J=self.J, d=self.d, krylov=krylov, res2=res2,
                                       jprec=self.jprec, stream=self.nt_stream)
        torch.cuda.synchronize()

    def replay_newton(self):
        self.disc.newton_graph_replay(self.nt_stream)


def autotune(disc, run, candidates=None, reps: int = 5, log=None):
    """Time run() under every (assembly path, threads, lds_bytes) launch
    shape, keep the fastest on disc and return (best, table of (path,
    threads, lds, ms)).  Path 0 evaluates each SCVF once into records, path 1
    is the fused kernel, path 2 is path 0 with dim+1 gather lanes per
    (node, element) (64 threads only), path 3 the plane march of structured
    hexahedral grids."""
    import torch

    if candidates is None:
        candidates = [(p, t, l) for p in (0, 1) for t in (64, 128, 256) for l in (16384, 32768, 65536)]
        candidates += [(2, 64, l) for l in (16384, 65536)]
        # the plane march (structured grids of parallelepipeds; elsewhere path
        # 0): the budget in KiB is its node planes per work unit
        candidates += [(3, 256, l) for l in (8192, 16384, 32768)]
    from . import NSFV1Error

    table = []
    for p, t, l in candidates:
        disc.set_assembly_path(p)
        disc.set_launch(t, l)
        try:
            run()  # builds the plan for this shape
        except NSFV1Error as err:
            # a shape whose budget cannot hold one node row of this mesh
            if err.name != "ARGUMENT":
                raise
            if log:
                log(f"launch path={p} threads={t} lds={l}: not possible ({err})")
            continue
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(reps):
            run()
        e.record()
        torch.cuda.synchronize()
        ms = s.elapsed_time(e) / reps
        table.append((p, t, l, ms))
        if log:
            log(f"launch path={p} threads={t} lds={l}: {ms:.3f} ms")
    if not table:
        raise ValueError("no launch shape fits this mesh")
    best = min(table, key=lambda r: r[3])
    disc.set_assembly_path(best[0])
    disc.set_launch(best[1], best[2])
    run()
    torch.cuda.synchronize()
    return best, table
