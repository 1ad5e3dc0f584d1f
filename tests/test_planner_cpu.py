"""Planner training and the numpy forward pass (CPU only)."""
import numpy as np


def test_train_learns_an_argmin():
    from amethyst_amd import planner
    rng = np.random.default_rng(0)
    n = 600
    X = rng.standard_normal((n, planner.NFEAT)).astype(np.float32)
    # synthetic times: plan 0 best when feature 0 > 0, plan 1 otherwise
    T = np.ones((n, 3)) * 2.0
    T[X[:, 0] > 0, 0] = 1.0
    T[X[:, 0] <= 0, 1] = 1.0
    w = planner.train(X, T, epochs=60)
    pred = planner.mlp_reference(w, X).argmin(axis=1)
    acc = np.mean(pred == T.argmin(axis=1))
    assert acc > 0.95, acc


def test_plan_space():
    from amethyst_amd import planner
    assert planner.NPLANS == 12
    assert planner.PLANS[planner.ANALYTIC][1] == "w+1"
    assert planner.plan_width("w+1", 11, 5) == 12
    assert planner.plan_width("est", 11, 5) == 5
    assert planner.plan_width("est+1", 11, 12) == 12
