"""Validation sets and evaluation metrics on the GPU (through the C ABI),
checked against the oracle's restatement of metric.rs.

Tolerances:
  * validation scores: bit-exact against binned prediction of the same
    trees (the same per-tree f64 additions in model order);
  * unweighted AUC, binary_error, multi_error, NDCG: bit-exact (integer or
    half-integer partial sums; NDCG repeats the reference's chains);
  * logloss metrics and weighted AUC: relative 1e-10 (a fixed parallel
    summation order, and device exp/log within an ulp of libm's).
"""
import ctypes as C

import numpy as np
import pytest

from tests._pkg import datagen, lgb, small_binary
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def _eval(b, data, metric, **kw):
    ids = lgb._capi.METRIC_IDS
    at = np.ascontiguousarray(kw.get("eval_at", (1, 2, 3, 4, 5)), np.int32)
    o = lgb._capi.MetricOpts(ids[metric], kw.get("top_k", 1), kw.get("threads", 1),
                             len(at) if metric == "ndcg" else 0, at.ctypes.data if metric == "ndcg" else None, 0,
                             None)
    out = np.zeros(max(len(at), 1), np.float64)
    cnt = C.c_int32(0)
    lgb._capi.check(lgb._capi.lib().lgbmhip_eval(b._ctx, data, C.byref(o), lgb.ptr(out), C.byref(cnt)), b._ctx)
    return out[: cnt.value]


def _with_scores(score, label, weight=None, params=None, group=None):
    """A booster whose validation set 1 holds `score` as its init scores
    (no tree trained), so the metric kernels see arbitrary inputs."""
    n = len(label)
    params = params or dict(objective="binary", num_leaves=7, max_bin=15)
    rng = np.random.default_rng(0)
    X = rng.normal(size=(max(n, 64), 3)).astype(np.float32)
    yt = (rng.random(len(X)) < 0.5).astype(np.float32)
    K = int(params.get("num_class", 1)) if params["objective"] == "multiclass" else 1
    if params["objective"] == "multiclass":
        yt = rng.integers(0, K, len(X)).astype(np.float32)
    gt = None
    if params["objective"] == "lambdarank":
        gt = [len(X) // 64] * 63 + [len(X) - 63 * (len(X) // 64)]
        yt = rng.integers(0, 3, len(X)).astype(np.float32)
    ds = lgb.Dataset(X, label=yt, group=gt, params=params)
    b = lgb.Booster(params, ds)
    Xv = rng.normal(size=(n, 3)).astype(np.float32)
    vb = lgb.Dataset(Xv, label=label, reference=ds, params=params).bins
    qb = None if group is None else np.ascontiguousarray(np.concatenate([[0], np.cumsum(group)]), np.int32)
    w = None if weight is None else np.ascontiguousarray(weight, np.float32)
    init = np.ascontiguousarray(score, np.float64).ravel()
    lab = np.ascontiguousarray(label, np.float32)
    idx = C.c_int32(0)
    lgb._capi.check(lgb._capi.lib().lgbmhip_add_valid(b._ctx, lgb.ptr(vb), n, lgb.ptr(lab), lgb.ptr(w), lgb.ptr(qb),
                                                      0 if group is None else len(group), lgb.ptr(init),
                                                      C.byref(idx)), b._ctx)
    assert idx.value == 1
    return b


@pytest.mark.parametrize("n,ties", [(1, False), (777, False), (5000, True), (300000, True)])
def test_auc_bit_exact(n, ties):
    rng = np.random.default_rng(n)
    s = rng.normal(size=n)
    if ties:
        s = np.round(s, 1)
        s[rng.random(n) < 0.1] = -0.0   # -0.0 and +0.0 are one score
        s[rng.random(n) < 0.1] = 0.0
    y = (rng.random(n) < 0.3).astype(np.float32)
    b = _with_scores(s, y)
    print(f"auc n={n} ties={ties}", flush=True)
    got = _eval(b, 1, "auc")[0]
    assert got == O.metric_auc(s, y)


def test_auc_single_class_is_one_and_weighted_within_tolerance():
    rng = np.random.default_rng(3)
    n = 20000
    s = np.round(rng.normal(size=n), 2)
    b = _with_scores(s, np.ones(n, np.float32))
    assert _eval(b, 1, "auc")[0] == 1.0
    y = (rng.random(n) < 0.4).astype(np.float32)
    w = (rng.random(n) * 2).astype(np.float32)
    b = _with_scores(s, y, weight=w)
    got = _eval(b, 1, "auc")[0]
    assert got == pytest.approx(O.metric_auc(s, y, w), rel=RTOL)


@pytest.mark.parametrize("weighted", [False, True])
def test_binary_pointwise_metrics(weighted):
    rng = np.random.default_rng(4)
    n = 100003
    s = rng.normal(size=n) * 3
    s[:50] = 40.0      # probabilities that clamp at the epsilon
    s[50:100] = -40.0
    y = (rng.random(n) < 0.5).astype(np.float32)
    w = (rng.random(n) + 0.5).astype(np.float32) if weighted else None
    b = _with_scores(s, y, weight=w)
    ll = _eval(b, 1, "binary_logloss")[0]
    assert ll == pytest.approx(O.metric_point("binary_logloss", s, y, w), rel=RTOL)
    err = _eval(b, 1, "binary_error")[0]
    ref = O.metric_point("binary_error", s, y, w)
    if weighted:
        assert err == pytest.approx(ref, rel=RTOL)
    else:
        assert err == ref


@pytest.mark.parametrize("top_k", [1, 2])
def test_multiclass_pointwise_metrics(top_k):
    rng = np.random.default_rng(5)
    n, K = 50000, 5
    s = rng.normal(size=(K, n))
    s[:, :100] = np.round(s[:, :100])   # ties between classes
    y = rng.integers(0, K, n).astype(np.float32)
    params = dict(objective="multiclass", num_class=K, num_leaves=7, max_bin=15)
    b = _with_scores(s, y, params=params)
    ll = _eval(b, 1, "multi_logloss")[0]
    assert ll == pytest.approx(O.metric_point("multi_logloss", s, y, K=K), rel=RTOL)
    err = _eval(b, 1, "multi_error", top_k=top_k)[0]
    assert err == O.metric_point("multi_error", s, y, K=K, top_k=top_k)


@pytest.mark.parametrize("threads", [1, 7])
def test_ndcg_bit_exact(threads):
    rng = np.random.default_rng(6)
    sizes = np.concatenate([[1, 2, 5000], rng.geometric(1 / 60, 700)])
    n = int(sizes.sum())
    rel = rng.integers(0, 5, n).astype(np.float32)
    rel[1:3] = 0   # a query without relevant documents counts 1
    s = np.round(rng.normal(size=n), 1)   # ties keep index order
    params = dict(objective="lambdarank", num_leaves=7, max_bin=15)
    b = _with_scores(s, rel, params=params, group=sizes)
    qb = np.concatenate([[0], np.cumsum(sizes)])
    for at in ((1, 2, 3, 4, 5), (1, 10, 30), (3, 1000)):
        got = _eval(b, 1, "ndcg", eval_at=at, threads=threads)
        ref = O.metric_ndcg(s, rel, qb, at, threads)
        assert got.tolist() == ref.tolist(), at


def _binary_sets(n_train=60000, n_valid=20000, seed=7):
    X, y = small_binary(n_train + n_valid, F=10, seed=seed)
    return X[:n_train], y[:n_train], X[n_train:], y[n_train:]


@pytest.mark.parametrize("quant", [False, True])
def test_valid_scores_equal_binned_prediction(quant):
    Xt, yt, Xv, yv = _binary_sets()
    params = dict(objective="binary", num_leaves=31, max_bin=63, use_quantized_grad=quant,
                  num_grad_quant_bins=4)
    ds = lgb.Dataset(Xt, label=yt, params=params)
    dv = lgb.Dataset(Xv, label=yv, reference=ds, params=params)
    b = lgb.Booster(params, ds)
    b.add_valid(dv, "valid")
    for _ in range(8):
        b.update()
    # a set attached after training started catches up tree by tree
    dv2 = lgb.Dataset(Xv[:7777], label=yv[:7777], reference=ds, params=params)
    b.add_valid(dv2, "late")
    for _ in range(4):
        b.update()
    pv = b.predict_binned(dv.bins)[:, 0]
    assert np.array_equal(b.valid_scores(0), pv)
    assert np.array_equal(b.valid_scores(1), b.predict_binned(dv2.bins)[:, 0])
    # training scores through the same evaluation entry point
    ts = b.train_scores()
    assert _eval(b, 0, "auc")[0] == O.metric_auc(ts, yt)
    assert _eval(b, 1, "auc")[0] == O.metric_auc(pv, yv)
    assert _eval(b, 1, "binary_logloss")[0] == pytest.approx(O.metric_point("binary_logloss", pv, yv), rel=RTOL)


def test_multiclass_valid_scores_and_metrics():
    rng = np.random.default_rng(8)
    n, K = 40000, 3
    X = rng.normal(size=(n, 6)).astype(np.float32)
    y = np.argmax(X[:, :K] + rng.gumbel(size=(n, K)), 1).astype(np.float32)
    params = dict(objective="multiclass", num_class=K, num_leaves=15, max_bin=63)
    ds = lgb.Dataset(X[:30000], label=y[:30000], params=params)
    dv = lgb.Dataset(X[30000:], label=y[30000:], reference=ds, params=params)
    b = lgb.Booster(params, ds)
    b.add_valid(dv, "valid")
    for _ in range(4):
        b.update()
    pv = b.predict_binned(dv.bins)   # [rows][K]
    vs = b.valid_scores(0).reshape(K, -1)
    assert np.array_equal(vs, pv.T)
    assert _eval(b, 1, "multi_error")[0] == O.metric_point("multi_error", vs, y[30000:], K=K)
    assert _eval(b, 1, "multi_logloss")[0] == pytest.approx(
        O.metric_point("multi_logloss", vs, y[30000:], K=K), rel=RTOL)


def test_lambdarank_ndcg_on_train_and_valid():
    X, rel, qb = datagen.rank_2m(600, 300)
    nq_t = 400
    cut = int(qb[nq_t])
    params = dict(objective="lambdarank", num_leaves=31, max_bin=63, eval_at=[1, 3, 10])
    ds = lgb.Dataset(X[:cut], label=rel[:cut], group=np.diff(qb[: nq_t + 1]), params=params)
    dv = lgb.Dataset(X[cut:], label=rel[cut:], group=np.diff(qb[nq_t:]), reference=ds, params=params)
    b = lgb.Booster(params, ds)
    b.add_valid(dv, "valid")
    for _ in range(3):
        b.update()
    res = b.eval_valid() + b.eval_train()
    vs = b.valid_scores(0)
    ref_v = O.metric_ndcg(vs, rel[cut:], qb[nq_t:] - cut, (1, 3, 10))
    ref_t = O.metric_ndcg(b.train_scores(), rel[:cut], qb[: nq_t + 1], (1, 3, 10))
    assert [r.metric_value for r in res] == ref_v.tolist() + ref_t.tolist()
    assert [r.metric_name for r in res[:3]] == ["ndcg@1", "ndcg@3", "ndcg@10"]


def test_train_with_early_stopping():
    Xt, yt, Xv, yv = _binary_sets(20000, 10000, seed=9)
    yt = np.where(np.random.default_rng(1).random(len(yt)) < 0.3, 1 - yt, yt)   # noisy labels overfit
    params = dict(objective="binary", num_leaves=63, max_bin=63, learning_rate=0.5, min_data_in_leaf=5,
                  metric="binary_logloss", early_stopping_round=5)
    ds = lgb.Dataset(Xt, label=yt, params=params)
    dv = lgb.Dataset(Xv, label=yv, reference=ds, params=params)
    hist = []
    b = lgb.train(params, ds, num_boost_round=200, valid_sets=[dv], valid_names=["v"],
                  callbacks=[lambda env: hist.append(env.evaluation_result_list[0].metric_value)])
    assert 1 <= b.best_iteration < len(hist) <= 200
    assert b.best_iteration == int(np.argmin(hist)) + 1
    assert b.best_score["v"]["binary_logloss"] == min(hist)
    assert len(hist) - b.best_iteration == 5
