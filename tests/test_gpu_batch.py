"""Batched class trees: the K trees of a multiclass iteration sharing every
growth launch (batch_trees=True, the default) against the same trees grown
on separate growth lanes (batch_trees=False), and against the oracle.

The batch changes which launches carry a tree, never what is computed for it:
trees, leaf values and training scores must be identical bit for bit, with
and without captured graphs, in both histogram forms, with categorical
features, bagging and validation sets.
"""
import numpy as np
import pytest

from tests._pkg import lgb
from tests._trees import compare_or_near_tie
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def _data(n=24000, F=10, K=5, seed=21):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, F)).astype(np.float32)
    X[:, F - 1] = rng.integers(0, 12, size=n).astype(np.float32)   # a categorical column
    W = rng.normal(size=(F, K))
    y = np.argmax(X @ W + rng.gumbel(size=(n, K)), 1).astype(np.float32)
    return X, y


def _train(X, y, params, iters, valid=None):
    ds = lgb.Dataset(X, label=y, params=params)
    b = lgb.Booster(params, ds)
    if valid is not None:
        b.add_valid(lgb.Dataset(valid[0], label=valid[1], reference=ds, params=params), "v")
    for _ in range(iters):
        b.update()
    return b


def _same_model(a, b):
    assert a.num_trees() == b.num_trees()
    for t in range(a.num_trees()):
        ta, tb = a.tree(t), b.tree(t)
        for f in ("split_feature", "threshold_in_bin", "threshold", "decision_type", "left_child", "right_child",
                  "leaf_value", "leaf_count"):
            np.testing.assert_array_equal(getattr(ta, f), getattr(tb, f), err_msg=f"tree {t} {f}")
    np.testing.assert_array_equal(a.train_scores(), b.train_scores())
    assert a.model_to_string() == b.model_to_string()


@pytest.mark.parametrize("quant,graph,batch", [(False, True, True), (True, True, True), (True, False, True),
                                               (True, True, 2), (False, False, 3), (True, True, 16)])
def test_batched_trees_equal_lane_trees(quant, graph, batch):
    X, y = _data()
    params = dict(objective="multiclass", num_class=5, num_leaves=31, max_bin=63, min_data_in_leaf=20,
                  categorical_feature=[9], use_graph=graph)
    if quant:
        params.update(use_quantized_grad=True, num_grad_quant_bins=16, stochastic_rounding=True,
                      quant_train_renew_leaf=True)
    print(f"[batch] quant={quant} graph={graph} batch={batch}", flush=True)
    a = _train(X, y, dict(params, batch_trees=batch), 6)
    b = _train(X, y, dict(params, batch_trees=False), 6)
    _same_model(a, b)
    a.free()
    b.free()


def test_batched_trees_with_bagging_and_a_validation_set():
    X, y = _data(seed=22)
    Xv, yv = _data(n=5000, seed=23)
    params = dict(objective="multiclass", num_class=5, num_leaves=15, max_bin=63, bagging_fraction=0.6,
                  bagging_freq=1, bagging_seed=7, metric="multi_logloss")
    a = _train(X, y, dict(params, batch_trees=True), 5, valid=(Xv, yv))
    b = _train(X, y, dict(params, batch_trees=False), 5, valid=(Xv, yv))
    _same_model(a, b)
    np.testing.assert_array_equal(a.eval_valid()[0].metric_value, b.eval_valid()[0].metric_value)


def test_batched_trees_match_oracle_on_the_same_gradients():
    X, y = _data(n=12000, K=4, seed=24)
    params = dict(objective="multiclass", num_class=4, num_leaves=63, max_bin=63, min_data_in_leaf=20,
                  categorical_feature=[9])
    ds = lgb.Dataset(X, label=y, params=params)
    b = lgb.Booster(dict(params, batch_trees=True), ds)
    ob = O.Booster(O.Dataset(X, y, params), params)
    n = len(y)
    for it in range(3):
        b.update()
        g, h = b.last_gradients()
        for k in range(4):
            ob.train_tree(g[k * n:(k + 1) * n], h[k * n:(k + 1) * n])
    for t in range(12):
        bias = 0.0
        if t < 4:
            p = max(1e-15, float(np.count_nonzero(y == t)) / n)
            bias = np.log(p)
        assert compare_or_near_tie(b.tree(t), ob.tree(t), 0.1, bias)


@pytest.mark.parametrize("quant", [False, True])
def test_batched_trees_search_from_histogram_units(quant):
    """Batched rounds whose searches sum the histogram units themselves
    (late_round 1: every round after the first) give the trees of lanes whose
    rounds all go through the slab reduction (late_round 0)."""
    X, y = _data(n=30000, seed=23)
    params = dict(objective="multiclass", num_class=5, num_leaves=63, max_bin=63, min_data_in_leaf=10,
                  categorical_feature=[9])
    if quant:
        params.update(use_quantized_grad=True, num_grad_quant_bins=16, stochastic_rounding=True,
                      quant_train_renew_leaf=True)
    boosters = []
    for batch, late in ((True, 1), (False, 0)):
        ds = lgb.Dataset(X, label=y, params=params)
        b = lgb.Booster(dict(params, batch_trees=batch), ds)
        b.set_tuning(late_round=late)
        for _ in range(5):
            b.update()
        boosters.append(b)
    _same_model(*boosters)
    for b in boosters:
        b.free()
