"""Writes tests/golden/fv1_golden.npz: oracle Jacobians and defects of tiny
seeded meshes (2x2 jittered quads/triangles, one hexahedron, one cube split
into 6 tetrahedra) for every upwind, steady and time-dependent.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from plugin_navierstokes_amd import meshes as M  # noqa: E402


def cases():
    meshes = {
        "quad": M.cavity_2d(2, M.QUAD, 0.1, 1),
        "tri": M.cavity_2d(2, M.TRI, 0.1, 1),
        "hex": M.cavity_3d(1, M.HEX),
        "tet": M.cavity_3d(1, M.TET),
    }
    for name, mesh in meshes.items():
        for up in ["no", "full", "skewed", "lps"]:
            for td in [False, True]:
                yield f"{name}_{up}_{'td' if td else 'st'}", mesh, up, td


def main():
    out = {}
    for key, mesh, up, td in cases():
        u = M.random_state(mesh, 3)
        uold = M.random_state(mesh, 4) if td else None
        dt = 0.05 if td else 0.0
        p = O.Params(dim=mesh.dim, upwind=up, nu=0.02)
        rp, ci = O.pattern(mesh.et, mesh.nv, mesh.conn)
        J, d = O.assemble(p, mesh.et, mesh.xyz, mesh.conn, u, uold, dt, s_a=1.0, s_m=1.0 if td else 0.0,
                          parts=3 if td else 1, rowptr=rp, colidx=ci)
        out[key + "/xyz"] = mesh.xyz
        out[key + "/conn"] = mesh.conn
        out[key + "/u"] = u
        if td:
            out[key + "/uold"] = uold
        out[key + "/rowptr"] = rp
        out[key + "/colidx"] = ci
        out[key + "/J"] = J
        out[key + "/d"] = d
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "fv1_golden.npz"), **out)


if __name__ == "__main__":
    main()
