"""Features of at most 16 bins stored as 4-bit fields (pack_u4, the default;
bin.rs:92-121 Dense4): the packed rows change where a bin is stored, never
what is computed from it. Trees, leaf values, scores and model text must
equal the byte layout's (pack_u4=False) bit for bit, in both histogram
forms, with categorical and missing-value features, device binning,
bagging, validation sets, rollback and batched class trees; the trees grown
on the same gradients equal the oracle's; integer histograms of packed rows
equal the CPU integer reference.
"""
import numpy as np
import pytest

from tests._pkg import lgb
from tests._trees import compare_or_near_tie
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def _data(n=30000, seed=31, classes=0):
    """6 normal columns (byte-wide bins at max_bin 63) and 11 columns of at
    most 16 bins: 7 of k in [2, 14] values, a 10-category column, a column
    with missing values, a binary one and a 3-valued one."""
    rng = np.random.default_rng(seed)
    cols = [rng.normal(size=n) for _ in range(6)]
    for k in (2, 3, 5, 7, 11, 13, 14):
        cols.append(np.floor(rng.random(n) * k))
    cols.append(rng.integers(0, 10, n).astype(np.float64))          # categorical
    miss = np.floor(rng.random(n) * 6)
    miss[rng.random(n) < 0.15] = np.nan                             # missing values
    cols.append(miss)
    cols.append((rng.random(n) < 0.3).astype(np.float64))
    cols.append(np.floor(rng.random(n) * 3) - 1.0)
    X = np.stack(cols, 1).astype(np.float32)
    z = X[:, 0] + 0.5 * X[:, 6] - 0.3 * X[:, 9] + 0.2 * np.nan_to_num(X[:, 14]) + 0.4 * (X[:, 13] % 3)
    if classes:
        W = rng.normal(size=(X.shape[1], classes))
        y = np.argmax(np.nan_to_num(X) @ W + rng.gumbel(size=(n, classes)), 1).astype(np.float32)
    else:
        y = (rng.random(n) < 1 / (1 + np.exp(-z))).astype(np.float32)
    return X, y


CAT = [13]
BASE = dict(objective="binary", num_leaves=31, max_bin=63, min_data_in_leaf=20, categorical_feature=CAT)
QUANT = dict(use_quantized_grad=True, num_grad_quant_bins=16, stochastic_rounding=True, quant_train_renew_leaf=True)


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


def test_packed_layout_counts_the_4bit_features():
    X, y = _data(4000)
    a = _train(X, y, dict(BASE, pack_u4=True), 0)
    b = _train(X, y, dict(BASE, pack_u4=False), 0)
    la, lb = a.bin_layout(), b.bin_layout()
    F = a.train_set.num_features
    n4 = sum(1 for f in range(F) if a.train_set.feature_meta(f)["num_bin"] <= 16)
    assert n4 >= 10, n4
    assert la == dict(row_bytes=(F - n4) + (n4 + 1) // 2, row_stride=la["row_stride"], u4_features=n4)
    assert lb["row_bytes"] == F and lb["u4_features"] == 0
    assert la["row_stride"] < lb["row_stride"]
    a.free()
    b.free()


@pytest.mark.parametrize("quant,graph", [(False, True), (True, True), (True, False)])
def test_packed_trees_equal_byte_trees(quant, graph):
    X, y = _data()
    params = dict(BASE, use_graph=graph, **(QUANT if quant else {}))
    print(f"[u4] quant={quant} graph={graph}", flush=True)
    a = _train(X, y, dict(params, pack_u4=True), 8)
    b = _train(X, y, dict(params, pack_u4=False), 8)
    _same_model(a, b)
    a.free()
    b.free()


@pytest.mark.parametrize("quant", [False, True])
def test_packed_trees_match_oracle_on_the_same_gradients(quant):
    X, y = _data(n=20000, seed=32)
    params = dict(BASE, num_leaves=63, **(QUANT if quant else {}))
    ds = lgb.Dataset(X, label=y, params=params)
    b = lgb.Booster(dict(params, pack_u4=True), ds)
    assert b.bin_layout()["u4_features"] >= 10
    ob = O.Booster(O.Dataset(X, y, params), params)
    n = len(y)
    p = min(max(float(np.count_nonzero(y > 0)) / n, 1e-15), 1.0 - 1e-15)
    for it in range(3):
        b.update()
        g, h = b.last_gradients()
        ob.train_tree(g, h)
        bias = float(np.log(p / (1.0 - p))) if it == 0 else 0.0
        if compare_or_near_tie(b.tree(it), ob.tree(it), 0.1, bias):
            np.testing.assert_array_equal(b.leaf_of_row(), ob.leaf_of_row())
    b.free()


def test_packed_device_binning_bagging_validation_and_rollback():
    X, y = _data(seed=33)
    Xv, yv = _data(n=6000, seed=34)
    params = dict(BASE, device_binning=True, bagging_fraction=0.7, bagging_freq=1, bagging_seed=5,
                  metric="binary_logloss", **QUANT)
    a = _train(X, y, dict(params, pack_u4=True), 6, valid=(Xv, yv))
    b = _train(X, y, dict(params, pack_u4=False), 6, valid=(Xv, yv))
    assert a.bin_layout()["u4_features"] >= 10
    _same_model(a, b)
    np.testing.assert_array_equal(a.eval_valid()[0].metric_value, b.eval_valid()[0].metric_value)
    # validation scores (walks over the packed rows) equal prediction of the same rows
    np.testing.assert_array_equal(a.valid_scores(0), a.predict(Xv, raw_score=True))
    # rollback subtracts the last tree over the packed rows (training and validation)
    a.rollback_one_iter()
    b.rollback_one_iter()
    _same_model(a, b)
    np.testing.assert_array_equal(a.valid_scores(0), b.valid_scores(0))
    a.free()
    b.free()


@pytest.mark.parametrize("quant", [False, True])
def test_packed_multiclass_batched_trees(quant):
    X, y = _data(n=20000, seed=35, classes=4)
    params = dict(BASE, objective="multiclass", num_class=4, num_leaves=31, **(QUANT if quant else {}))
    a = _train(X, y, dict(params, pack_u4=True, batch_trees=True), 4)
    b = _train(X, y, dict(params, pack_u4=False, batch_trees=False), 4)
    _same_model(a, b)
    a.free()
    b.free()


def test_all_features_4bit_at_max_bin_15():
    X, y = _data(seed=36)
    params = dict(BASE, max_bin=15, **QUANT)
    a = _train(X, y, dict(params, pack_u4=True), 6)
    b = _train(X, y, dict(params, pack_u4=False), 6)
    F = a.train_set.num_features
    assert a.bin_layout() == dict(row_bytes=(F + 1) // 2, row_stride=a.bin_layout()["row_stride"], u4_features=F)
    _same_model(a, b)
    a.free()
    b.free()


def test_packed_integer_histogram_bit_exact():
    X, y = _data(n=40000, seed=37)
    params = dict(BASE, pack_u4=True, use_quantized_grad=True, num_grad_quant_bins=16)
    ds = lgb.Dataset(X, label=y, params=params)
    b = lgb.Booster(params, ds)
    rng = np.random.default_rng(38)
    gq = rng.integers(-8, 9, len(y)).astype(np.int8)
    hq = rng.integers(0, 17, len(y)).astype(np.int8)
    rows = np.sort(rng.choice(len(y), 25000, replace=False)).astype(np.uint32)
    B = max(ds.feature_meta(f)["num_bin"] for f in range(ds.num_features))
    out = np.zeros(ds.num_features * B, np.int64)
    nb = np.zeros(1, np.int32)
    lib = lgb._capi.lib()
    lgb._capi.check(lib.lgbmhip_build_histogram_int(b._ctx, lgb.ptr(rows), len(rows), lgb.ptr(gq), lgb.ptr(hq),
                                                     lgb.ptr(out), lgb.ptr(nb)), b._ctx)
    mfb = np.array([ds.feature_meta(f)["most_freq_bin"] for f in range(ds.num_features)], np.int32)
    cpu = O.hist_int(ds.bins, mfb, B, rows, gq, hq)
    np.testing.assert_array_equal(out.reshape(ds.num_features, B), cpu)
    b.free()
