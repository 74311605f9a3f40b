"""Learned Newton initial guess for BDF time stepping (3D).

A per-node MLP (12 -> 128 -> 128 -> 4, ReLU, BF16 weights and activations,
FP32 accumulation on the MFMA units; csrc/nsfv1_mlp.hip) maps
    (u^n, u^n - u^{n-1}, 1-ring mean of u^n - u^n)      (velocity + pressure)
to a correction of the linear extrapolation:
    guess = 2 u^n - u^{n-1} + MLP(features) * out_scale.
It is trained on trajectories of the plain solver with the library's own
forward/backward/Adam kernels.  During time stepping the FP64 defect judges
every prediction: the step starts from whichever of {prediction, linear
extrapolation} has the smaller defect, so the method is never worse than
the plain one by more than the cost of the two defect evaluations.

No counterpart in the reference plugin.  Training:

    python -m plugin_navierstokes_amd.learned train --out weights.npz
"""
from __future__ import annotations

import json
import time

import numpy as np

from . import NavierStokesFV1, _check, _ptr, _stream_ptr, load_library

NPARAM = 20752
KIN, OUT = 16, 4


def _lib():
    return load_library()


class MLP:
    """Device parameters, BF16 copy and Adam state of the predictor network."""

    def __init__(self, params: np.ndarray | None = None, seed: int = 0):
        import torch

        self.dev = torch.device("cuda:0")
        if params is None:
            params = init_params(seed)
        params = np.ascontiguousarray(params, dtype=np.float32)
        if params.shape != (NPARAM,):
            raise ValueError(f"expected {NPARAM} parameters")
        self.P = torch.from_numpy(params.copy()).to(self.dev)
        self.WB = torch.empty(NPARAM, dtype=torch.int16, device=self.dev)
        self.M = torch.zeros_like(self.P)
        self.V = torch.zeros_like(self.P)
        self.G = torch.zeros_like(self.P)
        self.step = 0
        _check(_lib().nsfv1_mlp_pack(_ptr(self.P), _ptr(self.WB), None))
        self._work = None

    def forward(self, X, cache: bool = False):
        import torch

        n = X.shape[0]
        Y = torch.empty(n, OUT, dtype=torch.float32, device=self.dev)
        z1 = torch.empty(n, 128, dtype=torch.float32, device=self.dev) if cache else None
        z2 = torch.empty(n, 128, dtype=torch.float32, device=self.dev) if cache else None
        _check(_lib().nsfv1_mlp_forward(_ptr(self.P), _ptr(self.WB), _ptr(X), n, _ptr(Y), _ptr(z1), _ptr(z2), None))
        return (Y, z1, z2) if cache else Y

    def gradient(self, X, T):
        """Loss and gradient for a batch (the gradient lands in self.G)."""
        import ctypes as C

        import torch

        n = X.shape[0]
        Y, z1, z2 = self.forward(X, cache=True)
        dY = torch.empty_like(Y)
        nw = C.c_int64(0)
        _check(_lib().nsfv1_mlp_sizes(n, None, C.byref(nw)))
        if self._work is None or self._work.numel() < nw.value:
            self._work = torch.empty(nw.value, dtype=torch.float32, device=self.dev)
        loss = C.c_double(0.0)
        _check(_lib().nsfv1_mlp_mse(_ptr(Y), _ptr(T), n, _ptr(dY), _ptr(self._work), C.byref(loss), None))
        _check(_lib().nsfv1_mlp_backward(_ptr(self.WB), _ptr(X), _ptr(z1), _ptr(z2), _ptr(dY), n, _ptr(self.G),
                                         _ptr(self._work), None))
        return loss.value

    def adam_step(self, lr, b1=0.9, b2=0.999, eps=1e-8):
        self.step += 1
        _check(_lib().nsfv1_mlp_adam(_ptr(self.P), _ptr(self.M), _ptr(self.V), _ptr(self.G), self.step, float(lr),
                                     float(b1), float(b2), float(eps), _ptr(self.WB), None))

    def train_step(self, X, T, lr):
        loss = self.gradient(X, T)
        self.adam_step(lr)
        return loss

    def params(self) -> np.ndarray:
        return self.P.cpu().numpy().copy()


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform initialisation, zero padding, small output layer."""
    rng = np.random.default_rng(seed)
    H, IN = 128, 12
    P = np.zeros(NPARAM, dtype=np.float32)
    W1 = np.zeros((H, KIN), np.float32)
    W1[:, :IN] = rng.uniform(-1, 1, (H, IN)) * np.sqrt(6.0 / IN)
    W2 = (rng.uniform(-1, 1, (H, H)) * np.sqrt(6.0 / H)).astype(np.float32)
    W3 = np.zeros((16, H), np.float32)
    W3[:OUT] = rng.uniform(-1, 1, (OUT, H)) * np.sqrt(6.0 / H) * 0.1
    P[0:H * KIN] = W1.ravel()
    o2 = H * KIN + H
    P[o2:o2 + H * H] = W2.ravel()
    o3 = o2 + H * H + H
    P[o3:o3 + 16 * H] = W3.ravel()
    return P


class LearnedPredictor:
    """Features -> MLP -> guess on one discretisation's mesh."""

    def __init__(self, disc: NavierStokesFV1, mlp: MLP, in_scale, out_scale):
        import torch

        self.disc, self.mlp = disc, mlp
        dev = mlp.dev
        self.in_scale = torch.as_tensor(np.asarray(in_scale, np.float32), device=dev)
        self.out_scale = torch.as_tensor(np.asarray(out_scale, np.float32), device=dev)
        nv = disc.mesh.nv
        self.X = torch.empty(nv, KIN, dtype=torch.float32, device=dev)

    def features(self, u, u_prev):
        _check(_lib().nsfv1_predict_features(self.disc._h, _ptr(u), _ptr(u_prev), _ptr(self.in_scale), _ptr(self.X),
                                             None))
        return self.X

    def target(self, u_next, u, u_prev):
        import torch

        T = torch.empty(u.shape[0], OUT, dtype=torch.float32, device=self.mlp.dev)
        _check(_lib().nsfv1_predict_target(self.disc._h, _ptr(u_next), _ptr(u), _ptr(u_prev), _ptr(self.out_scale),
                                           _ptr(T), None))
        return T

    def guess(self, u, u_prev):
        import torch

        Y = self.mlp.forward(self.features(u, u_prev))
        g = torch.empty_like(u)
        _check(_lib().nsfv1_predict_guess(self.disc._h, _ptr(u), _ptr(u_prev), _ptr(Y), _ptr(self.out_scale), _ptr(g),
                                          None))
        return g


def save(path, params, in_scale, out_scale, meta: dict):
    np.savez(path, params=np.asarray(params, np.float32), in_scale=np.asarray(in_scale, np.float32),
             out_scale=np.asarray(out_scale, np.float32), meta=np.frombuffer(json.dumps(meta).encode(), np.uint8))


def load(path):
    z = np.load(path, allow_pickle=False)
    return z["params"], z["in_scale"], z["out_scale"], json.loads(bytes(z["meta"]).decode())


# ---------------------------------------------------------------------
# the channel workload: plain solver, trajectories, training
# ---------------------------------------------------------------------
def channel_case(seed: int, cells=(96, 48, 48), dt: float = 0.005, tune: bool = False, smoothing=(2, 2, 0.7)):
    """Seeded channel: Re in [100, 500], 1-4 obstacle boxes.  Returns
    (mesh, disc, newton, stepper, u0, info).  tune: pick the assembly path
    and launch shape of the fine level with the autotuner first;
    smoothing: (pre, post, omega) of the multigrid's block-Jacobi sweeps."""
    import torch

    from . import F32, PART_A, PART_M
    from . import meshes as M
    from .batch import autotune
    from .solver import BDFStepper, NewtonSolver, build_multigrid

    rng = np.random.default_rng(1000 + seed)
    re = float(rng.uniform(100.0, 500.0))
    nbox = int(rng.integers(1, 5))
    mesh = M.channel_3d(cells, n_boxes=nbox, seed=seed)
    disc = NavierStokesFV1("u,v,w,p", "Inner")
    disc.set_stabilization("fields")
    disc.set_upwind("lps")  # continuous in u: Newton reaches 1e-8 (full/skewed switch discontinuously)
    disc.set_kinematic_viscosity(1.0 / re)
    disc.set_mesh(mesh)
    dofs, vals = M.channel_dirichlet(mesh)
    build_multigrid(disc, mesh, M.channel_dirichlet, pre=smoothing[0], post=smoothing[1], omega=smoothing[2])
    nt = NewtonSolver(disc, dofs, vals, jprec=F32, rtol=1e-8, atol=0.0, maxit=30, lin_rtol=1e-3, lin_maxit=200,
                      restart=30, multigrid=True)
    stepper = BDFStepper(nt, dt, reference="old", mg_lag=True)
    u0 = M.channel_initial_state(mesh, (dofs, vals))
    info = {"seed": seed, "Re": re, "boxes": nbox}
    if tune:
        ud = torch.from_numpy(u0).to("cuda:0")
        best, _ = autotune(disc, lambda: disc.assemble(ud, ud, dt, s_m=1.0, parts=PART_A | PART_M, jprec=F32),
                           reps=3)
        info["launch"] = {"path": best[0], "threads": best[1], "lds_bytes": best[2], "ms": best[3]}
    return mesh, disc, nt, stepper, u0, info


def run_trajectory(case, nsteps, predictor=None, record=None, log=None, warm=0):
    """BDF1 start, `warm` plain BDF2 steps (not returned), then nsteps BDF2
    steps (with the predictor's candidates when given).  record(u_next, u,
    u_prev) is called for every BDF2 step; returns per returned step
    (seconds, newton its, choice, initial residual / reference)."""
    import torch

    mesh, disc, nt, stepper, u0, info = case
    u = torch.from_numpy(u0).to("cuda:0")
    u, _ = stepper.step(u)
    torch.cuda.synchronize()
    out = []
    for k in range(warm + nsteps):
        u_prev = stepper.u_prev
        t0 = time.perf_counter()
        if predictor is not None and k >= warm:
            cand = [predictor.guess(u, u_prev), stepper.extrapolate(u)]
            u_new, st = stepper.step(u, candidates=cand)
            choice = stepper.last_choice
        else:
            u_new, st = stepper.step(u, guess=stepper.extrapolate(u))
            choice = -1
        torch.cuda.synchronize()
        dt_s = time.perf_counter() - t0
        if not st.converged:
            raise RuntimeError(f"step {k} did not converge: {st.residuals}")
        if record is not None:
            record(u_new, u, u_prev)
        if k >= warm:
            out.append((dt_s, st.iterations, choice, st.residuals[0] / st.reference))
        if log:
            log(f"step {k}: {dt_s:.3f} s newton {st.iterations} gmres {st.linear_iterations} choice {choice} "
                f"initial residual {st.residuals[0] / st.reference:.2e}")
        u = u_new
    return out


def train(out_path, n_traj=16, steps=24, skip=2, epochs=12, batch=65536, lr=1e-3, seed=0, cells=(96, 48, 48),
          log=print):
    """Generate trajectories with the plain solver, fit the MLP with the
    library's kernels, save weights and scales."""
    import torch

    t_start = time.perf_counter()
    feats, targs = [], []
    in_scale = out_scale = None
    mlp = MLP(seed=seed)
    for tr in range(n_traj):
        case = channel_case(tr, cells)
        pred = LearnedPredictor(case[1], mlp, np.ones(12), np.ones(4))
        rec = []

        def record(u_next, u, u_prev):
            rec.append((u_next.clone(), u.clone(), u_prev.clone()))

        run_trajectory(case, steps, record=record)
        # scales from the first trajectory: unit-variance features and targets
        if in_scale is None:
            # features() fills one buffer: each step's slice is copied
            Xs = torch.cat([pred.features(u, up)[:, :12].clone() for _, u, up in rec[skip:]])
            in_scale = (1.0 / Xs.std(0).clamp_min(1e-12)).cpu().numpy()
            Ts = torch.cat([pred.target(un, u, up) for un, u, up in rec[skip:]])
            out_scale = Ts.std(0).clamp_min(1e-12).cpu().numpy()
            # the pressure keeps its linear extrapolation: a per-node network
            # cannot see the global pressure coupling
            out_scale[3] = 0.0
        pred = LearnedPredictor(case[1], mlp, in_scale, out_scale)
        tscale = np.where(out_scale > 0, out_scale, 1.0)
        pred_t = LearnedPredictor(case[1], mlp, in_scale, tscale)
        for un, u, up in rec[skip:]:
            feats.append(pred.features(u, up).clone())
            t = pred_t.target(un, u, up)
            t[:, 3] = 0.0
            targs.append(t)
        log(f"trajectory {tr} ({case[5]}): {len(rec)} steps, {time.perf_counter() - t_start:.0f} s")
        del case, pred, rec
        torch.cuda.empty_cache()
    X = torch.cat(feats)
    T = torch.cat(targs)
    del feats, targs
    n = X.shape[0]
    g = torch.Generator(device="cuda:0").manual_seed(seed)
    hist = []
    for ep in range(epochs):
        perm = torch.randperm(n, device="cuda:0", generator=g)
        tot, cnt = 0.0, 0
        lr_ep = lr * (0.5 ** (ep // 4))
        for b0 in range(0, n - batch + 1, batch):
            idx = perm[b0:b0 + batch]
            tot += mlp.train_step(X[idx].contiguous(), T[idx].contiguous(), lr_ep)
            cnt += 1
        hist.append(tot / max(cnt, 1))
        log(f"epoch {ep}: loss {hist[-1]:.4f} (target variance 1), {time.perf_counter() - t_start:.0f} s")
    meta = {"trajectories": n_traj, "steps": steps, "skip": skip, "samples": int(n), "epochs": epochs,
            "batch": batch, "lr": lr, "loss": hist, "cells": list(cells), "upwind": "lps", "dt": 0.005}
    save(out_path, mlp.params(), in_scale, out_scale, meta)
    return meta


def compare(weights, seeds=(100, 101), steps=8, cells=(96, 48, 48), log=print, warm=6):
    """Plain vs learned s/step on held-out channels (same states, same
    1e-8 criterion).  Returns a dict of means."""
    params, in_scale, out_scale, meta = load(weights)
    mlp = MLP(params)
    res = {"plain": [], "learned": [], "plain_newton": [], "learned_newton": [], "learned_choice": []}
    for sd in seeds:
        for mode in ("plain", "learned"):
            # 1 + 1 block-Jacobi sweeps at omega 0.9: the fastest measured on
            # this channel (undamped sweeps take 1.5x as long here)
            case = channel_case(sd, cells, tune=True, smoothing=(1, 1, 0.9))
            res["launch"] = case[5].get("launch")
            res["mesh_sizes"] = {"nv": int(case[0].nv), "ne": int(case[0].ne), "nnz": int(case[1].nnz())}
            pred = LearnedPredictor(case[1], mlp, in_scale, out_scale) if mode == "learned" else None
            out = run_trajectory(case, steps, predictor=pred, log=None, warm=warm)
            res[mode] += [o[0] for o in out]
            res[mode + "_newton"] += [o[1] for o in out]
            res[mode + "_r0"] = res.get(mode + "_r0", []) + [o[3] for o in out]
            if mode == "learned":
                res["learned_choice"] += [o[2] for o in out]
            log(f"seed {sd} {mode}: {np.mean([o[0] for o in out]):.3f} s/step, "
                f"newton {np.mean([o[1] for o in out]):.2f}, initial residual {np.mean([o[3] for o in out]):.2e}")
    return {"plain_s_per_step": float(np.mean(res["plain"])), "learned_s_per_step": float(np.mean(res["learned"])),
            "plain_newton_per_step": float(np.mean(res["plain_newton"])),
            "learned_newton_per_step": float(np.mean(res["learned_newton"])),
            "prediction_used_fraction": float(np.mean(np.asarray(res["learned_choice"]) == 0)),
            "plain_initial_residual": float(np.mean(res["plain_r0"])),
            "learned_initial_residual": float(np.mean(res["learned_r0"])),
            "steps_timed": len(res["plain"]), "warm_steps": warm, "launch": res.get("launch"),
            "mesh_sizes": res.get("mesh_sizes")}


if __name__ == "__main__":
    import argparse
    import sys

    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd")
    t = sub.add_parser("train")
    t.add_argument("--out", required=True)
    t.add_argument("--trajectories", type=int, default=16)
    t.add_argument("--steps", type=int, default=24)
    t.add_argument("--epochs", type=int, default=12)
    t.add_argument("--cells", type=int, nargs=3, default=[96, 48, 48])
    c = sub.add_parser("compare")
    c.add_argument("--weights", required=True)
    c.add_argument("--steps", type=int, default=8)
    c.add_argument("--cells", type=int, nargs=3, default=[96, 48, 48])
    a = ap.parse_args()

    def _log(m):
        print(m, flush=True)

    if a.cmd == "train":
        m = train(a.out, a.trajectories, a.steps, epochs=a.epochs, cells=tuple(a.cells), log=_log)
        print(json.dumps({k: v for k, v in m.items() if k != "loss"}), flush=True)
    elif a.cmd == "compare":
        print(json.dumps(compare(a.weights, steps=a.steps, cells=tuple(a.cells), log=_log)), flush=True)
    else:
        ap.print_help()
        sys.exit(2)
