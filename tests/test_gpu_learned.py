"""The trained predictor on a held-out channel (seed, Re and obstacles not
seen in training): the FP64 defect judges each prediction, so a step never
starts from a guess worse than the linear extrapolation, and the network
lowers the defect of the initial guess on average."""
from __future__ import annotations

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "plugin_navierstokes_amd", "data",
                       "predictor_channel.npz")


def test_trained_predictor_gain(cuda):
    from plugin_navierstokes_amd.learned import MLP, LearnedPredictor, channel_case, load, run_trajectory

    params, in_scale, out_scale, meta = load(WEIGHTS)
    mlp = MLP(params)
    plain = run_trajectory(channel_case(100), 4, warm=6)
    case = channel_case(100)
    learned = run_trajectory(case, 4, predictor=LearnedPredictor(case[1], mlp, in_scale, out_scale), warm=6)
    r_plain = np.array([o[3] for o in plain])
    r_learned = np.array([o[3] for o in learned])
    print(f"initial residual / reference: plain {r_plain}, learned {r_learned}; "
          f"choices {[o[2] for o in learned]}; newton plain {[o[1] for o in plain]} learned {[o[1] for o in learned]}")
    # both runs follow the same trajectory (converged to 1e-8), so step k's
    # extrapolation is the same guess in both: the judged start is never worse
    assert all(o[2] in (0, 1) for o in learned)
    assert np.all(r_learned <= r_plain * (1 + 1e-4))
    assert r_learned.mean() < r_plain.mean()
