"""The trained predictor on a held-out channel (seed, Re and obstacles not
seen in training), over 24 steps after 20 plain warm-up steps:
- the FP64 defect judges each prediction against the linear extrapolation,
  so no step starts from a guess worse than the plain method's;
- the learned method takes fewer Newton iterations in total than the plain
  one (linear extrapolation) to the same 1e-8 criterion;
- the judge against the old state (no extra defect evaluation) falls back
  only where the prediction's defect exceeds the old state's."""
from __future__ import annotations

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "plugin_navierstokes_amd", "data",
                       "predictor_channel.npz")
WARM, STEPS = 20, 24


def _run(mode, mlp=None, scales=None):
    from plugin_navierstokes_amd.learned import LearnedPredictor, channel_case, run_trajectory

    case = channel_case(100)
    pred = LearnedPredictor(case[1], mlp, *scales) if mlp is not None else None
    return run_trajectory(case, STEPS, predictor=pred, warm=WARM, judge=mode)


def test_trained_predictor_gain(cuda):
    from plugin_navierstokes_amd.learned import MLP, load

    params, in_scale, out_scale, meta = load(WEIGHTS)
    mlp = MLP(params)
    plain = _run("linear")
    learned = _run("linear", mlp, (in_scale, out_scale))
    r_plain = np.array([o[3] for o in plain])
    r_learned = np.array([o[3] for o in learned])
    n_plain = np.array([o[1] for o in plain])
    n_learned = np.array([o[1] for o in learned])
    t_plain = np.mean([o[0] for o in plain])
    t_learned = np.mean([o[0] for o in learned])
    print(f"newton per step: plain {n_plain.mean():.3f}, learned {n_learned.mean():.3f}; "
          f"s/step plain {t_plain:.4f}, learned {t_learned:.4f}; choices {[o[2] for o in learned]}")
    print(f"initial residual / reference: plain {np.median(r_plain):.2e}, learned {np.median(r_learned):.2e} (medians)")
    assert len(plain) == len(learned) == STEPS
    # both runs follow the same trajectory (converged to 1e-8), so step k's
    # linear extrapolation is the same guess in both: the judged start is
    # never worse than the plain one
    assert all(o[2] in (0, 1) for o in learned)
    assert np.all(r_learned <= r_plain * (1 + 1e-3))
    assert n_learned.sum() < n_plain.sum()


def test_old_state_judge(cuda):
    from plugin_navierstokes_amd.learned import MLP, load

    params, in_scale, out_scale, meta = load(WEIGHTS)
    out = _run("old", MLP(params), (in_scale, out_scale))
    assert len(out) == STEPS and all(o[2] in (0, 1) for o in out)
    # a prediction that passes starts the step: its defect is below the old
    # state's (relative initial residual < 1)
    assert all(o[3] <= 1.0 for o in out if o[2] == 1)
