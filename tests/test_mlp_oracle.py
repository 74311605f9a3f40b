"""Pins the FP64 MLP oracle (the checker of the learned-guess kernels):
bf16 rounding on known values, gradients by central finite differences of
the unrounded network, Adam's first step in closed form, features by brute
force on a small mesh, and the library's parameter layout."""
from __future__ import annotations

import numpy as np

import oracle.mlp_oracle as MO
from plugin_navierstokes_amd import meshes as M


def test_bf16_rounding_known_values():
    one = 1.0
    ulp = 2.0 ** -7  # bf16 spacing at 1
    vals = np.array([one, one + ulp / 2, one + ulp / 2 + 2 ** -20, one + 3 * ulp / 2, -2.5, 0.0, 1e-30],
                    dtype=np.float32)
    want = np.array([one, one, one + ulp, one + 2 * ulp, -2.5, 0.0], dtype=np.float64)
    got = MO.bf16(vals)
    assert np.array_equal(got[:6], want)  # ties to even, above-tie up
    assert abs(got[6] - 1e-30) < 1e-32


def test_layout_matches_library_constants():
    assert MO.NPARAM == 20752
    assert (MO.OFF_B1, MO.OFF_W2, MO.OFF_B2, MO.OFF_W3, MO.OFF_B3) == (2048, 2176, 18560, 18688, 20736)


def test_gradient_finite_differences():
    rng = np.random.default_rng(0)
    P = MO.init_params(1).astype(np.float64)
    P[MO.OFF_B1:MO.OFF_W2] = rng.normal(0, 0.1, MO.H)
    P[MO.OFF_B2:MO.OFF_W3] = rng.normal(0, 0.1, MO.H)
    X = np.zeros((7, MO.KIN))
    X[:, :12] = rng.normal(size=(7, 12))
    T = rng.normal(size=(7, 4))
    _, g, _ = MO.loss_and_grad(P, X, T, rounding=False)
    idx = np.concatenate([rng.choice(np.arange(MO.OFF_W1, MO.OFF_B1).reshape(MO.H, MO.KIN)[:, :12].ravel(), 6),
                          rng.integers(MO.OFF_B1, MO.OFF_W2, 3), rng.integers(MO.OFF_W2, MO.OFF_B2, 6),
                          rng.integers(MO.OFF_B2, MO.OFF_W3, 3),
                          rng.choice(np.arange(MO.OFF_W3, MO.OFF_W3 + 4 * MO.H), 6), MO.OFF_B3 + np.arange(4)])
    for p in idx:
        h = 1e-6
        Pp, Pm = P.copy(), P.copy()
        Pp[p] += h
        Pm[p] -= h
        fd = (MO.loss_and_grad(Pp, X, T, rounding=False)[0] - MO.loss_and_grad(Pm, X, T, rounding=False)[0]) / (2 * h)
        assert abs(fd - g[p]) <= 1e-6 * max(1.0, abs(fd)), (p, fd, g[p])


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(3)
    P = MO.init_params(2).astype(np.float64)
    G = rng.normal(size=MO.NPARAM)
    P1, M1, V1 = MO.adam(P.copy(), np.zeros(MO.NPARAM), np.zeros(MO.NPARAM), G, 1, 1e-3, eps=0.0)
    # first bias-corrected step moves every non-padding weight by exactly lr * sign(g)
    live = (P1 != 0) & (P != 0)
    np.testing.assert_allclose((P - P1)[live], (1e-3 * np.sign(G))[live], rtol=1e-9)
    assert np.all(P1[MO.OFF_B3 + 4:] == 0) and np.all(P1[MO.OFF_W3 + 4 * MO.H:MO.OFF_B3] == 0)


def test_features_bruteforce():
    mesh = M.cavity_3d(2, M.TET)
    rng = np.random.default_rng(5)
    u = rng.normal(size=(mesh.nv, 4))
    up = rng.normal(size=(mesh.nv, 4))
    scale = rng.uniform(0.5, 2, 12)
    import oracle as O

    rp, ci = O.pattern(O.TET, mesh.nv, mesh.conn)
    X = MO.features(rp, ci, u, up, scale)
    # brute force: neighbours = nodes sharing an element
    for v in [0, 13, 26]:
        nb = sorted(set(mesh.conn[np.any(mesh.conn == v, axis=1)].ravel()) - {v})
        want = np.concatenate([u[v], u[v] - up[v], u[nb].mean(0) - u[v]]) * scale
        np.testing.assert_allclose(X[v, :12], want, rtol=1e-14)
        assert np.all(X[v, 12:] == 0)
