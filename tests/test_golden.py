"""The oracle reproduces its committed golden vectors (tests/golden/, made by
tests/golden/make_golden.py) bit for bit; the GPU path matches them within
the FP64 tolerance."""
import os

import numpy as np
import pytest

import oracle as O
from plugin_navierstokes_amd import meshes as M
from tests.golden.make_golden import cases

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "fv1_golden.npz"))
KEYS = [c[0] for c in cases()]


def _case(key):
    for k, mesh, up, td in cases():
        if k == key:
            return mesh, up, td
    raise KeyError(key)


@pytest.mark.parametrize("key", KEYS)
def test_oracle_reproduces_golden(key):
    mesh, up, td = _case(key)
    p = O.Params(dim=mesh.dim, upwind=up, nu=0.02)
    J, d = O.assemble(p, mesh.et, G[key + "/xyz"], G[key + "/conn"], G[key + "/u"],
                      G[key + "/uold"] if td else None, 0.05 if td else 0.0, s_a=1.0, s_m=1.0 if td else 0.0,
                      parts=3 if td else 1, rowptr=G[key + "/rowptr"], colidx=G[key + "/colidx"])
    assert np.array_equal(J, G[key + "/J"]) and np.array_equal(d, G[key + "/d"])


@pytest.mark.gpu
@pytest.mark.parametrize("key", KEYS)
def test_gpu_matches_golden(cuda, key):
    import torch

    from plugin_navierstokes_amd import F64, PART_A, PART_M, NavierStokesFV1

    mesh, up, td = _case(key)
    disc = NavierStokesFV1("u,v,p" if mesh.dim == 2 else "u,v,w,p", "Inner")
    disc.set_stabilization("fields")
    disc.set_upwind(up)
    disc.set_kinematic_viscosity(0.02)
    disc.set_mesh(et=mesh.et, xyz=G[key + "/xyz"], conn=G[key + "/conn"])
    r, c = disc.pattern()
    assert np.array_equal(r, G[key + "/rowptr"]) and np.array_equal(c, G[key + "/colidx"])
    u = torch.from_numpy(G[key + "/u"]).to(cuda)
    uo = torch.from_numpy(G[key + "/uold"]).to(cuda) if td else None
    J, d = disc.assemble(u, uo, 0.05 if td else 0.0, s_a=1.0, s_m=1.0 if td else 0.0,
                         parts=(PART_A | PART_M) if td else PART_A, jprec=F64)
    Jr, dr = G[key + "/J"], G[key + "/d"]
    p = O.Params(dim=mesh.dim, upwind=up, nu=0.02)
    Ja, da = O.assemble_abs(p, mesh.et, G[key + "/xyz"], G[key + "/conn"], G[key + "/u"],
                            G[key + "/uold"] if td else None, 0.05 if td else 0.0, s_a=1.0, s_m=1.0 if td else 0.0,
                            parts=3 if td else 1, rowptr=r, colidx=c)
    J, d = J.cpu().numpy(), d.cpu().numpy()
    assert np.all(np.abs(J - Jr) <= 1e-12 * np.abs(Jr) + 1e-13 * Ja)
    assert np.all(np.abs(d - dr) <= 1e-12 * np.abs(dr) + 1e-13 * da)
