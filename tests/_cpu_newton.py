"""CPU fixpoint-Newton solver on top of the oracle assembly (test
infrastructure): Dirichlet rows set to identity with defect u - g, the
linear systems solved directly with scipy.  Used to check the oracle at the
solution level against the reference's lid-driven-cavity tables and as the
reference of the GPU Newton tests."""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O


def dirichlet_dofs(mesh, lid_speed=1.0, pin_pressure=True):
    """Velocity Dirichlet DoFs of walls and (regularised) lid, plus one pressure DoF."""
    nc = mesh.dim + 1
    nodes = np.concatenate([mesh.boundary["wall"], mesh.boundary["lid"]])
    dofs = (nodes[:, None] * nc + np.arange(mesh.dim)[None, :]).ravel()
    vals = np.zeros(dofs.shape)
    lid = mesh.boundary["lid"]
    x = mesh.xyz[lid]
    inner = np.ones(len(lid), dtype=bool)
    for c in range(mesh.dim - 1):
        inner &= (x[:, c] > 1e-12) & (x[:, c] < 1 - 1e-12)
    lid_inner = set((lid[inner] * nc).tolist())
    vals = np.array([lid_speed if d in lid_inner else 0.0 for d in dofs])
    if pin_pressure:
        dofs = np.append(dofs, mesh.dim)  # pressure of node 0
        vals = np.append(vals, 0.0)
    order = np.argsort(dofs)
    return dofs[order].astype(np.int64), vals[order]


def apply_dirichlet(A: sp.csr_matrix, d: np.ndarray, u: np.ndarray, dofs, vals):
    A = A.tolil()
    for r in dofs:
        A.rows[r] = [r]
        A.data[r] = [1.0]
    A = A.tocsr()
    d = d.copy()
    d[dofs] = u[dofs] - vals
    return A, d


def newton(p: O.Params, mesh, u0=None, dofs=None, vals=None, rtol=1e-10, maxit=200, uold=None, dt=0.0,
           s_m=0.0, const=None):
    """Solve d(u) = 0 with the fixpoint Jacobian; returns (u, history)."""
    et = mesh.et
    nc = mesh.dim + 1
    rowptr, colidx = O.pattern(et, mesh.nv, mesh.conn)
    n = mesh.nv * nc
    u = np.zeros(n) if u0 is None else u0.ravel().copy()
    if dofs is not None:
        u[dofs] = vals
    hist = []
    parts = 1 | (2 if s_m else 0)
    for it in range(maxit):
        J, d = O.assemble(p, et, mesh.xyz, mesh.conn, u.reshape(-1, nc), uold, dt, s_a=1.0, s_m=s_m, parts=parts,
                          rowptr=rowptr, colidx=colidx)
        if const is not None:
            d = d + const
        A = sp.csr_matrix((J, colidx, rowptr), shape=(n, n))
        if dofs is not None:
            A, d = apply_dirichlet(A, d, u, dofs, vals)
        r = np.linalg.norm(d)
        hist.append(r)
        if r <= rtol * max(hist[0], 1e-300) or r < 1e-14:
            break
        u = u - spla.spsolve(A.tocsc(), d)
    return u.reshape(-1, nc), hist
