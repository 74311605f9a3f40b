"""Newton loop and BDF time stepping around the GPU assembly.

The reference plugin hands its element callbacks to ugcore, whose
DirichletBoundary, NewtonSolver and multistep time schemes drive them (none
of which are part of the plugin).  This module restates that driver for the
device-resident path:

* Dirichlet rows: J row -> identity, d -> u - g (ugcore DirichletBoundary).
* Newton: d(u) assembled in FP64; the Jacobian (fixpoint linearisation of
  navier_stokes_fv1.cpp:250-595, optionally FP32) is only the
  preconditioner of the correction, so it can be stored in FP32 without
  changing the converged FP64 solution; the correction is solved with GMRES
  + point-block Jacobi to a loose relative tolerance; u <- u - delta; the
  convergence test is ||d(u)|| <= rtol ||d(u0)|| in FP64.
* Time: BDF1 for the first step, BDF2 afterwards, mass part weights over
  three levels.  The stabilisation uses the two-point series (u^{n+1}, u^n,
  dt), the form navier_stokes_fv1.cpp:268-280 accepts (that code rejects
  more than two time points, so BDF2 with FV1 stabilisation is an extension
  of the reference).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import F32, F64, PART_A, PART_M, UPWIND_FRESH, UPWIND_RECORD, UPWIND_REPLAY


@dataclass
class NewtonStats:
    iterations: int = 0
    linear_iterations: int = 0
    residuals: list = field(default_factory=list)
    converged: bool = False
    seconds: float = 0.0
    reference: float = 0.0  # the norm the relative tolerance applies to
    fresh_residual: float | None = None  # with frozen upwind decisions: final defect with fresh decisions


def build_multigrid(disc, mesh, dirichlet, max_coarse_dofs=1200, pre=2, post=2, omega=0.7):
    """Geometric multigrid hierarchy for a structured mesh: coarsen until the
    coarsest level has at most max_coarse_dofs DoFs (or the cell counts stop
    being even), create a discretisation per coarse level (configuration is
    taken from disc), set its Dirichlet rows from dirichlet(mesh) -> (dofs,
    vals), and register the hierarchy with disc.  Returns the coarse discs."""
    from . import NavierStokesFV1
    from .meshes import coarsen

    coarse, prol, inj = [], [], []
    m = mesh
    while m.nv * (m.dim + 1) > max_coarse_dofs:
        if m.recipe is None or any(c % 2 for c in m.shape) or min(m.shape) < 2:
            break
        c, P, I = coarsen(m)
        d = NavierStokesFV1(",".join(["u", "v", "w"][: m.dim] + ["p"]), "Inner", dim=m.dim)
        d.set_mesh(c)
        dofs, _ = dirichlet(c)
        d.set_dirichlet(np.sort(np.asarray(dofs, dtype=np.int64)))
        coarse.append(d)
        prol.append(P)
        inj.append(I)
        m = c
    if not coarse:
        raise ValueError("mesh too small or not coarsenable for multigrid")
    if m.nv * (m.dim + 1) > max_coarse_dofs:
        raise ValueError(f"coarsest level has {m.nv * (m.dim + 1)} DoFs > {max_coarse_dofs}")
    disc.mg_set_hierarchy(coarse, prol, inj, pre=pre, post=post, omega=omega)
    disc.set_preconditioner("gmg")
    return coarse


class NewtonSolver:
    """Fixpoint/Newton iteration for d(u) = 0 with device buffers.

    With multigrid=True the correction equation is preconditioned by a
    geometric multigrid V-cycle (the disc needs a hierarchy from
    build_multigrid); otherwise by point-block Jacobi."""

    def __init__(self, disc, dirichlet_dofs, dirichlet_vals, jprec=F32, rtol=1e-8, atol=1e-14, maxit=60,
                 lin_rtol=1e-3, lin_maxit=400, restart=30, multigrid=False, freeze_upwind_after=None):
        import torch

        self.multigrid = multigrid
        # skewed upwind: decide anew for the first k-1 iterations, record at
        # iteration k-1, replay afterwards (None: always decide anew)
        self.freeze_after = freeze_upwind_after
        self.disc = disc
        self.jprec = jprec
        self.rtol, self.atol, self.maxit = rtol, atol, maxit
        self.lin_rtol, self.lin_maxit, self.restart = lin_rtol, lin_maxit, restart
        dofs = np.asarray(dirichlet_dofs, dtype=np.int64)
        order = np.argsort(dofs)
        self.dofs = dofs[order]
        disc.set_dirichlet(self.dofs)
        dev = torch.device("cuda:0")
        self.dev = dev
        self.g = torch.as_tensor(np.asarray(dirichlet_vals, dtype=np.float64)[order], device=dev)
        self.dofs_t = torch.as_tensor(self.dofs, device=dev)
        n = disc.num_dofs()
        self.J = torch.empty(disc.nnz(), dtype=torch.float32 if jprec == F32 else torch.float64, device=dev)
        self.d = torch.empty(n, dtype=torch.float64, device=dev)
        self.delta = torch.empty(n, dtype=torch.float64, device=dev)

    def impose(self, u):
        """Set the Dirichlet values in u (flat view)."""
        u.view(-1)[self.dofs_t] = self.g

    def defect(self, u, u_old=None, dt=0.0, s_m=0.0, const=None, want_J=True, mg_setup=True):
        parts = PART_A | (PART_M if s_m else 0)
        self.disc.assemble(u, u_old, dt, s_a=1.0, s_m=s_m, parts=parts, jprec=self.jprec,
                           J=self.J if want_J else None, d=self.d, want_J=want_J, check=False)
        if const is not None:
            self.d += const
        self.disc.apply_dirichlet(self.J if want_J else None, self.jprec, self.d, u, self.g)
        if want_J and self.multigrid and mg_setup:
            self.disc.mg_setup(self.jprec, u, u_old, dt, s_a=1.0, s_m=s_m, parts=parts)
        return self.d

    def defect_norm(self, u, u_old=None, dt=0.0, s_m=0.0, const=None) -> float:
        """||d(u)|| in FP64 (defect only, no Jacobian)."""
        import torch

        return float(torch.linalg.vector_norm(self.defect(u, u_old, dt, s_m, const, want_J=False)))

    def solve(self, u, u_old=None, dt=0.0, s_m=0.0, const=None, ref_norm=None, mg_lag=False,
              callback=None, assembled=None) -> NewtonStats:
        """Iterate until ||d(u)|| <= max(rtol * ref, atol), ref = ref_norm or
        the defect norm of the initial guess.  mg_lag keeps the multigrid
        levels of the first iteration for the whole solve (the Krylov
        operator is always the current Jacobian).  callback(it, u), if
        given, sees every iterate before its defect is assembled.
        assembled: ||d(u)|| when self.J, self.d (and the multigrid levels)
        already hold the full assembly at this u; the first iteration then
        uses them instead of assembling again."""
        import torch

        st = NewtonStats()
        t0 = time.perf_counter()
        self.impose(u)
        uf = u.view(-1)
        try:
            self._iterate(st, u, uf, u_old, dt, s_m, const, ref_norm, mg_lag, callback, assembled)
        finally:
            if self.freeze_after is not None:
                self.disc.set_upwind_decisions(UPWIND_FRESH)
        if self.freeze_after is not None and st.iterations >= self.freeze_after:
            st.fresh_residual = self.defect_norm(u, u_old, dt, s_m, const)
        self.disc.sync_errors()
        torch.cuda.synchronize()
        st.seconds = time.perf_counter() - t0
        return st

    def _iterate(self, st, u, uf, u_old, dt, s_m, const, ref_norm, mg_lag, callback=None, assembled=None):
        import torch

        k = self.freeze_after
        if assembled is not None and k is not None:
            raise ValueError("a reused first assembly cannot be combined with frozen upwind decisions")
        for it in range(self.maxit + 1):
            if callback is not None:
                callback(it, u)
            if k is not None:
                self.disc.set_upwind_decisions(UPWIND_FRESH if it < k - 1 else
                                               (UPWIND_RECORD if it == k - 1 else UPWIND_REPLAY))
            if it == 0 and assembled is not None:
                d, r = self.d, float(assembled)
            else:
                d = self.defect(u, u_old, dt, s_m, const, mg_setup=(it == 0 or not mg_lag))
                r = float(torch.linalg.vector_norm(d))
            st.residuals.append(r)
            ref = ref_norm if ref_norm is not None else st.residuals[0]
            st.reference = ref
            if r <= max(self.rtol * ref, self.atol):
                st.converged = True
                break
            if it == self.maxit:
                break
            its, _ = self.disc.linear_solve(self.J, d, self.delta, jprec=self.jprec, rtol=self.lin_rtol,
                                            maxit=self.lin_maxit, restart=self.restart)
            st.linear_iterations += its
            uf -= self.delta
            st.iterations += 1


class BDFStepper:
    """BDF1 start, then BDF2 with constant dt, for M du/dt + A(u) = 0.

    reference="old" measures convergence against the defect of the step's
    system at the old state u^n, so the stopping point does not depend on
    the initial guess (the guess only changes how many iterations it takes);
    reference="guess" uses the defect of the initial guess."""

    def __init__(self, newton: NewtonSolver, dt: float, reference: str = "guess", mg_lag: bool = False):
        self.nt = newton
        self.dt = dt
        self.u_prev = None  # u^{n-1}
        self.steps = 0
        self.reference = reference
        self.mg_lag = mg_lag
        self.last_choice = None

    def mass_const(self, w, scale):
        """scale * M w as a device vector (defect of the mass part, linear)."""
        import torch

        c = torch.empty_like(self.nt.d)
        self.nt.disc.assemble(w, None, 0.0, s_a=0.0, s_m=scale, parts=PART_M, jprec=self.nt.jprec, J=None, d=c,
                              want_J=False, check=False)
        return c

    def step(self, u_n, guess=None, candidates=None, reuse_last=False, fallback=None):
        """Returns (u^{n+1}, NewtonStats); u_n is not modified.  candidates:
        optional list of initial guesses; the one with the smallest FP64
        defect starts the Newton iteration (index kept in last_choice; a tie
        goes to the earlier one).  reuse_last: the last candidate is judged
        with the full assembly (Jacobian, defect, multigrid levels) instead
        of the defect alone; when it wins, Newton's first iteration reuses
        that assembly, so judging it costs no extra assembly.  fallback
        (with guess): the guess is judged against the old state instead, by
        its full assembly (reused when it passes): the step starts from the
        fallback when ||d(guess)|| exceeds ||d(u_n)|| of the step's system
        (last_choice 0), else from the guess (last_choice 1)."""
        import torch

        dt = self.dt
        if self.u_prev is None:
            s_m, const = 1.0 / dt, self.mass_const(u_n, -1.0 / dt)
        else:
            w = -2.0 * u_n + 0.5 * self.u_prev
            s_m, const = 1.5 / dt, self.mass_const(w, 1.0 / dt)
        ref = None
        if self.reference == "old":
            ref = self.nt.defect_norm(u_n, u_n, dt, s_m, const)
        assembled = None
        if candidates:
            reuse = reuse_last and self.nt.freeze_after is None
            norms = []
            for i, c in enumerate(candidates):
                full = reuse and i == len(candidates) - 1
                self.nt.impose(c)
                d = self.nt.defect(c, u_n, dt, s_m, const, want_J=full, mg_setup=full)
                norms.append(float(torch.linalg.vector_norm(d)))
            k = int(np.argmin(norms))
            self.last_choice = k
            u = candidates[k].clone()
            if reuse and k == len(candidates) - 1:
                assembled = norms[k]
        elif guess is not None and fallback is not None:
            if ref is None:
                ref = self.nt.defect_norm(u_n, u_n, dt, s_m, const)
            full = self.nt.freeze_after is None
            u = guess.clone()
            self.nt.impose(u)
            r = float(torch.linalg.vector_norm(self.nt.defect(u, u_n, dt, s_m, const, want_J=full, mg_setup=full)))
            if r <= ref:
                self.last_choice = 1
                assembled = r if full else None
            else:
                self.last_choice = 0
                u = fallback.clone()
            if self.reference != "old":
                ref = None
        else:
            u = (guess if guess is not None else u_n).clone()
        st = self.nt.solve(u, u_old=u_n, dt=dt, s_m=s_m, const=const, ref_norm=ref, mg_lag=self.mg_lag,
                           assembled=assembled)
        self.u_prev = u_n.clone()
        self.steps += 1
        return u, st

    def extrapolate(self, u_n):
        """Plain initial guess: linear extrapolation 2 u^n - u^{n-1}."""
        if self.u_prev is None:
            return u_n.clone()
        return 2.0 * u_n - self.u_prev


def cavity_dirichlet(mesh, lid_speed=1.0, pin_pressure=True):
    """Walls u = 0, regularised lid u = (lid_speed, 0[, 0]) away from the lid's
    edges, pressure pinned at node 0."""
    nc = mesh.dim + 1
    nodes = np.concatenate([mesh.boundary["wall"], mesh.boundary["lid"]])
    dofs = (nodes[:, None] * nc + np.arange(mesh.dim)[None, :]).ravel()
    lid = mesh.boundary["lid"]
    x = mesh.xyz[lid]
    inner = np.ones(len(lid), dtype=bool)
    for c in range(mesh.dim - 1):
        inner &= (x[:, c] > 1e-12) & (x[:, c] < 1 - 1e-12)
    lid_dofs = set((lid[inner] * nc).tolist())
    vals = np.array([lid_speed if d in lid_dofs else 0.0 for d in dofs.tolist()])
    if pin_pressure:
        dofs = np.append(dofs, mesh.dim)
        vals = np.append(vals, 0.0)
    unused = mesh.boundary.get("unused", np.zeros(0, dtype=np.int64))
    if len(unused):
        extra = (np.asarray(unused)[:, None] * nc + np.arange(nc)[None, :]).ravel()
        dofs = np.append(dofs, extra)
        vals = np.append(vals, np.zeros(len(extra)))
    return dofs.astype(np.int64), vals
