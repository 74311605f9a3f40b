"""Shared helpers of the parity tests: seeded cases, GPU assembly through the
C ABI, oracle assembly, and the stated tolerances.

Tolerance rule (FP64): |a - b| <= 1e-12 |b| + 1e-13 S, where S is the sum of
the absolute element contributions to the entry (the oracle's
``assemble_abs``); this covers rounding of sums with cancellation.
FP32 Jacobian: |a - b| <= 2^-17 S + 2^-20 |b| against the FP64 oracle.
"""
from __future__ import annotations

import numpy as np

import oracle as O
from plugin_navierstokes_amd import meshes as M

FP64_REL, FP64_ABS = 1e-12, 1e-13
FP32_ABS, FP32_REL = 2.0 ** -17, 2.0 ** -20


def small_mesh(et: int, n: int = 4, seed: int = 3, jitter: float = 0.1):
    if et in (M.TRI, M.QUAD):
        return M.cavity_2d(n, et, jitter, seed)
    return M.cavity_3d(max(2, n // 2 + 1) if et == M.HEX else max(2, n // 2), et, jitter, seed)


def oracle_params(dim, cfg: dict) -> O.Params:
    return O.Params(dim=dim, upwind=cfg.get("upwind", "full"), stab=cfg.get("stab", "fields"),
                    stab_upwind=cfg.get("stab_upwind"), pac=cfg.get("pac", False),
                    difflen=cfg.get("difflen", "raw"), stokes=cfg.get("stokes", False),
                    laplace=cfg.get("laplace", False), peclet=cfg.get("peclet", False),
                    exact=cfg.get("exact", 0.0), nu=cfg.get("nu", 0.05), rho=cfg.get("rho", 1.0),
                    source=cfg.get("source"))


def run_case(cuda, et, cfg, mesh=None, seed=5, time_dep=False, parts=1, s=(1.0, 0.0, 0.0), jprec=64, path=None):
    """GPU assembly through the C ABI and the oracle's on the same seeded
    state: (J, d, J_oracle, d_oracle, J_scale, d_scale, disc)."""
    import torch
    from plugin_navierstokes_amd import NavierStokesFV1, PART_A, PART_RHS

    mesh = mesh or small_mesh(et)
    dim = mesh.dim
    u = M.random_state(mesh, seed)
    uold = M.random_state(mesh, seed + 100) if time_dep else None
    dt = 0.01 if time_dep else 0.0
    disc = NavierStokesFV1("u,v,p" if dim == 2 else "u,v,w,p", "Inner")
    configure_disc(disc, cfg)
    disc.set_mesh(mesh)
    if path is not None:
        disc.set_launch(64, 16384)
        disc.set_assembly_path(path)
    ut = torch.from_numpy(u).to(cuda)
    uot = torch.from_numpy(uold).to(cuda) if time_dep else None
    J, d = disc.assemble(ut, uot, dt, s_a=s[0], s_m=s[1], s_f=s[2], parts=parts, jprec=jprec)
    rowptr, colidx = disc.pattern()
    p = oracle_params(dim, cfg)
    Jr, dr = O.assemble(p, et, mesh.xyz, mesh.conn, u, uold, dt, s_a=s[0], s_m=s[1], s_f=s[2], parts=parts,
                        rowptr=rowptr, colidx=colidx)
    Ja, da = O.assemble_abs(p, et, mesh.xyz, mesh.conn, u, uold, dt, s_a=s[0], s_m=s[1], parts=parts,
                            rowptr=rowptr, colidx=colidx)
    if parts & PART_RHS and cfg.get("source") is not None:
        da = da + np.abs(dr)
    return J.cpu().numpy(), d.cpu().numpy(), Jr, dr, Ja, da, disc


def configure_disc(disc, cfg: dict):
    """Apply a case configuration through the reference-style setters."""
    if cfg.get("stab", "fields") in ("fields", "flow"):
        disc.set_stabilization(cfg.get("stab", "fields"), cfg.get("difflen", "raw"))
    else:
        disc.set_stabilization_without()
    if cfg.get("upwind", "full") is not None:
        disc.set_upwind(cfg.get("upwind", "full"))
    if cfg.get("stab_upwind") is not None:
        disc.set_stabilization_upwind(cfg["stab_upwind"])
    elif cfg.get("stab", "fields") != "fields":
        disc.set_stabilization_upwind(cfg.get("upwind", "full"))
    if cfg.get("pac"):
        disc.set_pac_upwind(True)
    disc.set_kinematic_viscosity(cfg.get("nu", 0.05))
    disc.set_density(cfg.get("rho", 1.0))
    if cfg.get("source") is not None:
        disc.set_source(list(cfg["source"]))
    disc.set_stokes(cfg.get("stokes", False))
    disc.set_laplace(cfg.get("laplace", False))
    disc.set_peclet_blend(cfg.get("peclet", False))
    disc.set_exact_jacobian(cfg.get("exact", 0.0))


def fp64_violation(a, b, scale):
    """max of |a-b| / (1e-12|b| + 1e-13 S); <= 1 passes."""
    a, b, scale = np.asarray(a), np.asarray(b), np.asarray(scale)
    tol = FP64_REL * np.abs(b) + FP64_ABS * scale + 1e-300
    return float(np.max(np.abs(a - b) / tol)) if a.size else 0.0


def fp32_violation(a, b, scale):
    a, b, scale = np.asarray(a, dtype=np.float64), np.asarray(b), np.asarray(scale)
    tol = FP32_ABS * scale + FP32_REL * np.abs(b) + 1e-300
    return float(np.max(np.abs(a - b) / tol)) if a.size else 0.0


def max_rel_err(a, b, scale):
    a, b, scale = np.asarray(a, dtype=np.float64), np.asarray(b), np.asarray(scale)
    den = np.maximum(np.abs(b), 1e-30 + 0 * b)
    m = scale > 0
    return float(np.max(np.abs(a - b)[m] / np.maximum(den[m], 1e-3 * scale[m]))) if m.any() else 0.0


def worst(a, b, scale, rel=FP64_REL, ab=FP64_ABS):
    """Description of the worst entry (for assertion messages)."""
    a, b, scale = np.asarray(a, dtype=np.float64), np.asarray(b), np.asarray(scale)
    tol = rel * np.abs(b) + ab * scale + 1e-300
    i = int(np.argmax(np.abs(a - b) / tol))
    return f"idx={i} got={a[i]!r} want={b[i]!r} scale={scale[i]!r} viol={abs(a[i]-b[i])/tol[i]:.3g}"
