"""GPU parity of the data exports (navier_stokes_fv1.cpp:894-1491) through
the C ABI against the CPU oracle: every export at SCVF ips, corners and
local points, on all element types, with the stabilised ip velocity under
several stabilisation/upwind settings (steady and time-dependent).

Tolerance (FP64): |a - b| <= 1e-12 |b| + 1e-12 max|b| over the array."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from plugin_navierstokes_amd import NavierStokesFV1, NSFV1Error
from plugin_navierstokes_amd import meshes as M
from tests._parity import configure_disc, oracle_params, small_mesh

pytestmark = pytest.mark.gpu

ETS = [M.TRI, M.QUAD, M.TET, M.HEX, M.PRISM]
WHAT = ["velocity", "velocity_grad", "velocity_ip", "pressure", "pressure_grad"]


def _disc(mesh, cfg):
    disc = NavierStokesFV1("u,v,p" if mesh.dim == 2 else "u,v,w,p", "Inner")
    configure_disc(disc, cfg)
    disc.set_mesh(mesh)
    return disc


def _check(got, want, label):
    got = got.cpu().numpy()
    assert got.shape == want.shape, (got.shape, want.shape)
    tol = 1e-12 * np.abs(want) + 1e-12 * max(np.abs(want).max(), 1e-300)
    v = np.max(np.abs(got - want) / tol)
    print(f"{label}: violation {v:.3f}", flush=True)
    assert v <= 1.0


@pytest.mark.parametrize("at", ["scvf", "scv", "local"])
@pytest.mark.parametrize("what", WHAT)
@pytest.mark.parametrize("et", ETS)
def test_export_parity(cuda, et, what, at):
    import torch

    mesh = small_mesh(et, 5, seed=4, jitter=0.12)
    cfg = {"upwind": "full"}
    disc = _disc(mesh, cfg)
    u = M.random_state(mesh, 31)
    rng = np.random.default_rng(et)
    lips = None
    if at == "local":
        lips = rng.uniform(0.02, 0.3, size=(7, mesh.dim))
    got = disc.export(torch.from_numpy(u).to(cuda), what, at, lips)
    want = O.export(oracle_params(mesh.dim, cfg), et, mesh.xyz, mesh.conn, u, what, at, lips)
    _check(got, want, f"et={et} {what} at {at}")


@pytest.mark.parametrize("cfg", [{"upwind": "full"}, {"upwind": "lps", "difflen": "fivepoint"},
                                 {"upwind": "skewed", "difflen": "cor"}, {"upwind": "no", "source": (0.3, -0.2, 0.1)},
                                 {"upwind": "full", "stokes": True}, {"stab": "without", "upwind": "full"},
                                 {"upwind": "full", "stab_upwind": "lps"}],
                         ids=["full", "lps-5pt", "skewed-cor", "no-source", "stokes", "without", "full+lps"])
@pytest.mark.parametrize("timedep", [False, True], ids=["steady", "tdep"])
@pytest.mark.parametrize("et", ETS)
def test_stabilised_ip_velocity(cuda, et, cfg, timedep):
    import torch

    cfg = dict(cfg)
    if cfg.get("source") is not None:
        cfg["source"] = cfg["source"][: 2 if et in (M.TRI, M.QUAD) else 3]
    mesh = small_mesh(et, 5, seed=6, jitter=0.12)
    disc = _disc(mesh, cfg)
    u = M.random_state(mesh, 41)
    uold = M.random_state(mesh, 42) if timedep else None
    dt = 0.02 if timedep else 0.0
    got = disc.export(torch.from_numpy(u).to(cuda), "velocity_ip", "scvf",
                      u_old=None if uold is None else torch.from_numpy(uold).to(cuda), dt=dt)
    want = O.export(oracle_params(mesh.dim, cfg), et, mesh.xyz, mesh.conn, u, "velocity_ip", "scvf", None, uold, dt)
    _check(got, want, f"et={et} cfg={cfg} timedep={timedep}")


def test_export_errors(cuda):
    import torch

    mesh = small_mesh(M.QUAD, 3)
    disc = NavierStokesFV1("u,v,p", "Inner")
    disc.set_mesh(mesh)
    u = torch.zeros(mesh.nv, 3, dtype=torch.float64, device=cuda)
    with pytest.raises(NSFV1Error) as e:
        disc.export(u, "velocity_ip")
    assert e.value.name == "NO_STABILIZATION"
    with pytest.raises(NSFV1Error):
        disc.export(u, "velocity", "local", np.zeros((33, 2)))
    # plain exports need no stabilisation
    assert disc.export(u, "pressure", "scv").shape == (mesh.ne, 4, 1)
