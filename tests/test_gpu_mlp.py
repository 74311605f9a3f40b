"""GPU kernels of the learned guess against the FP64 oracle: MFMA forward,
backward gradients, Adam, features/guess/target, and a short fit."""
from __future__ import annotations

import numpy as np
import pytest

import oracle.mlp_oracle as MO
from plugin_navierstokes_amd import meshes as M

pytestmark = pytest.mark.gpu


def _params(seed):
    rng = np.random.default_rng(seed)
    P = MO.init_params(seed)
    P[MO.OFF_B1:MO.OFF_W2] = rng.normal(0, 0.2, MO.H)
    P[MO.OFF_B2:MO.OFF_W3] = rng.normal(0, 0.2, MO.H)
    P[MO.OFF_B3:MO.OFF_B3 + 4] = rng.normal(0, 0.2, 4)
    return P.astype(np.float32)


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    X = np.zeros((n, MO.KIN), np.float32)
    X[:, :MO.IN] = rng.normal(size=(n, MO.IN))
    return X


@pytest.mark.parametrize("n", [1, 16, 100, 4099])
def test_forward_matches_oracle(cuda, n):
    import torch

    from plugin_navierstokes_amd.learned import MLP

    P, X = _params(1), _inputs(n, 2)
    mlp = MLP(P)
    Y, z1, z2 = mlp.forward(torch.from_numpy(X).to(cuda), cache=True)
    Y = Y.cpu().numpy().astype(np.float64)
    ref, cache = MO.forward(P, X)
    S = MO.output_scale(P, X)
    err = np.abs(Y - ref)
    print(f"n={n}: max err/scale {np.max(err / S):.2e}, median {np.median(err / S):.2e}")
    assert np.all(err <= 2.0 ** -7 * S)
    assert np.median(err / S) <= 1e-5
    # pre-activations agree to FP32 accumulation accuracy
    z1s = np.abs(cache["x"]) @ np.abs(cache["W1"]).T + 1e-30
    assert np.all(np.abs(z1.cpu().numpy() - cache["z1"]) <= 1e-5 * z1s + 1e-6)


def test_backward_matches_oracle(cuda):
    import torch

    from plugin_navierstokes_amd.learned import MLP

    n = 3000
    P, X = _params(3), _inputs(n, 4)
    T = np.random.default_rng(5).normal(size=(n, 4)).astype(np.float32)
    mlp = MLP(P)
    loss = mlp.gradient(torch.from_numpy(X).to(cuda), torch.from_numpy(T).to(cuda))
    g = mlp.G.cpu().numpy().astype(np.float64)
    lref, gref, _ = MO.loss_and_grad(P, X, T)
    assert abs(loss - lref) <= 1e-3 * lref
    for name, a, b in [("W1", MO.OFF_W1, MO.OFF_B1), ("b1", MO.OFF_B1, MO.OFF_W2), ("W2", MO.OFF_W2, MO.OFF_B2),
                       ("b2", MO.OFF_B2, MO.OFF_W3), ("W3", MO.OFF_W3, MO.OFF_B3), ("b3", MO.OFF_B3, MO.NPARAM)]:
        rel = np.linalg.norm(g[a:b] - gref[a:b]) / np.linalg.norm(gref[a:b])
        print(f"grad {name}: relative error {rel:.2e}")
        assert rel <= 2e-3, name
    # padding carries no gradient
    assert np.all(g[MO.OFF_W1:MO.OFF_B1].reshape(MO.H, MO.KIN)[:, MO.IN:] == 0)
    assert np.all(g[MO.OFF_W3 + 4 * 128:MO.OFF_B3] == 0) and np.all(g[MO.OFF_B3 + 4:] == 0)


def test_adam_matches_oracle(cuda):
    import torch

    from plugin_navierstokes_amd.learned import MLP

    P = _params(6)
    rng = np.random.default_rng(7)
    mlp = MLP(P)
    Pr, Mr, Vr = P.astype(np.float64), np.zeros(MO.NPARAM), np.zeros(MO.NPARAM)
    for step in range(1, 4):
        G = rng.normal(size=MO.NPARAM).astype(np.float32)
        mlp.G.copy_(torch.from_numpy(G))
        mlp.adam_step(1e-2)
        Pr, Mr, Vr = MO.adam(Pr, Mr, Vr, G.astype(np.float64), step, 1e-2)
    np.testing.assert_allclose(mlp.P.cpu().numpy(), Pr, rtol=1e-5, atol=1e-6)
    wb = mlp.WB.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
    np.testing.assert_array_equal(wb.view(np.float32).astype(np.float64), MO.bf16(mlp.P.cpu().numpy()))


def test_features_guess_target(cuda):
    import torch

    import oracle as O
    from plugin_navierstokes_amd import NavierStokesFV1
    from plugin_navierstokes_amd.learned import MLP, LearnedPredictor

    mesh = M.channel_3d((8, 4, 4), n_boxes=1, seed=2)
    disc = NavierStokesFV1("u,v,w,p", "Inner")
    disc.set_mesh(mesh)
    rng = np.random.default_rng(8)
    u, u1, u2, u3, un = (rng.normal(size=(mesh.nv, 4)) for _ in range(5))
    scale = rng.uniform(0.5, 2, 24).astype(np.float32)
    osc = rng.uniform(0.5, 2, 4).astype(np.float32)
    osc[3] = 0.0  # a component without correction
    mlp = MLP(_params(9))
    pr = LearnedPredictor(disc, mlp, scale, osc)
    dev = lambda a: torch.from_numpy(a).to(cuda)
    du, du1, du2, du3 = dev(u), dev(u1), dev(u2), dev(u3)
    X, sig = pr.features(du, du1, du2, du3)
    X = X.cpu().numpy()
    rp, ci = O.pattern(O.TET, mesh.nv, mesh.conn)
    Xr, sr = MO.features(rp, ci, u, u1, u2, u3, scale)
    s = float(sig)
    assert abs(s - sr) <= 1e-13 * sr
    np.testing.assert_allclose(X, Xr, rtol=2e-7, atol=1e-6)
    T = pr.target(dev(un), du, du1, du2, sig).cpu().numpy()
    np.testing.assert_allclose(T, MO.target(un, u, u1, u2, osc, s), rtol=1e-6, atol=1e-6)
    g = pr.guess(du, du1, du2, du3).cpu().numpy()
    Y = mlp.forward(torch.from_numpy(X).to(cuda)).cpu().numpy().astype(np.float64)
    np.testing.assert_allclose(g, MO.guess(u, u1, u2, Y, osc, s), rtol=1e-12, atol=1e-12)
    # clamped features: a state far outside the scale
    Xc, _ = pr.features(du * 1e3, du1, du2, du3)
    assert float(Xc.abs().max()) == MO.CLAMP
    # a stationary state: sigma at its floor, features of the differences 0
    Xs, ss = pr.features(du, du, du, du)
    assert float(ss) == 1e-30 and float(Xs[:, 4:16].abs().max()) == 0.0


def test_fit_reduces_loss(cuda):
    """A smooth per-node target is learned by the library's own kernels."""
    import torch

    from plugin_navierstokes_amd.learned import MLP

    n = 65536
    X = _inputs(n, 10)
    T = np.stack([np.sin(X[:, 0]) * X[:, 1], X[:, 2] ** 2 - 1, np.tanh(X[:, 3] + X[:, 4]), 0.5 * X[:, 5]], 1)
    Xt, Tt = torch.from_numpy(X).to(cuda), torch.from_numpy(T.astype(np.float32)).to(cuda)
    mlp = MLP(seed=0)
    first = mlp.train_step(Xt, Tt, 3e-3)
    for _ in range(300):
        last = mlp.train_step(Xt, Tt, 3e-3)
    print(f"loss {first:.4f} -> {last:.4f}")
    assert last < 0.2 * first
