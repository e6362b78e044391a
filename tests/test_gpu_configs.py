"""The benchmark configurations themselves against the oracle, at their real
sizes (or, for the multiclass one, its first 200k rows):

- 10.5M x 28 binary (20 N(0,1) + 8 lognormal columns, 5% label flips),
  255 leaves: the plain FP64 form at max_bin 63 and 255 and the quantized
  form at max_bin 255. Bins equal the oracle's; the first two boosting
  trees equal the oracle's trees grown on the same gradients (structure,
  thresholds and leaf counts exactly, leaf values within rtol 1e-9), and
  every row lands in the same leaf. A differing split is accepted only as a
  near-tie: the two trees agree up to that split and its two gains are equal
  at f32 resolution (reported).
- 2M x 100 multiclass (10 classes, 127 leaves, 8 Zipf categorical columns
  declared): the first 200k rows, two iterations, all 20 trees.
- Prediction of a 1000-tree lambdarank model (255 leaves, 136 features) on
  1M rows, raw f32 and u8 bins: the oracle reads the same model text and its
  raw prediction of a 100k-row sample equals both forms bit for bit
  (predict.rs:231-339, tree.rs:605-652).
- The prediction window and early stopping (predict.rs:155-202) against a
  numpy loop over the leaf indices of the same model, and the oracle's.
"""
import numpy as np
import pytest

from tests._pkg import datagen, lgb
from tests._trees import compare_or_near_tie
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def data_10p5m():
    print("[configs] generating 10.5M x 28", flush=True)
    return datagen.binary_10p5m()


_ODS = {}


def _oracle_dataset(X, y, params, key):
    if key not in _ODS:
        print(f"[configs] oracle binning {key}", flush=True)
        _ODS.clear()
        _ODS[key] = O.Dataset(X, y, params)
    return _ODS[key]


def _init_scores(y, objective, K):
    """boost_from_score of the first iteration: binary.rs:129-154 (log odds of
    the positive share), multiclass.rs:156-158 (log of each class's share)."""
    n = len(y)
    if objective == "binary":
        p = min(max(float(np.count_nonzero(y > 0)) / n, 1e-15), 1.0 - 1e-15)
        return [np.log(p / (1.0 - p))]
    return [np.log(max(1e-15, float(np.count_nonzero(y == k)) / n)) for k in range(K)]


def _boost_same_gradients(b, ob, n_iter, K=1, lr=0.1, inits=None):
    """n_iter GPU iterations; after each, the oracle grows each class's tree on
    the GPU's gradients of that iteration. Returns the number of equal trees."""
    equal = 0
    n = b.train_set.num_data
    for it in range(n_iter):
        b.update()
        g, h = b.last_gradients()
        for k in range(K):
            ob.train_tree(g[k * n:(k + 1) * n], h[k * n:(k + 1) * n])
            t = it * K + k
            bias = inits[k] if (inits is not None and it == 0) else 0.0
            same = compare_or_near_tie(b.tree(t), ob.tree(t), lr, bias)
            if same and K == 1:
                np.testing.assert_array_equal(b.leaf_of_row(), ob.leaf_of_row())
            equal += same
    return equal


@pytest.mark.parametrize("max_bin,quant", [(63, False), (255, False), (255, True)])
def test_binary_10p5m_trees_match_oracle(data_10p5m, max_bin, quant):
    X, y = data_10p5m
    params = dict(objective="binary", num_leaves=255, max_bin=max_bin, min_data_in_leaf=20, learning_rate=0.1)
    if quant:
        params.update(use_quantized_grad=True, num_grad_quant_bins=16, stochastic_rounding=True,
                      quant_train_renew_leaf=True)
    print(f"[configs] binary 10.5M max_bin={max_bin} quant={quant}", flush=True)
    ds = lgb.Dataset(X, label=y, params=params)
    ods = _oracle_dataset(X, y, params, ("binary_10p5m", max_bin))
    # bins bit-exact against the oracle's host binning
    assert ds.num_features == ods.num_features == 28
    np.testing.assert_array_equal(ds.bins, ods.bins())
    b = lgb.Booster(params, ds)
    ob = O.Booster(ods, params)
    equal = _boost_same_gradients(b, ob, 2, inits=_init_scores(y, "binary", 1))
    assert b.tree(0).num_leaves == 255
    print(f"[configs] {equal} of 2 trees equal", flush=True)
    # training scores equal raw prediction of the training rows (end_to_end.rs:53-57), on a sample
    idx = np.arange(0, len(y), 97)
    raw = b.predict(X[idx], raw_score=True)
    np.testing.assert_array_equal(raw, b.train_scores()[idx])
    b.free()


def test_multiclass_2m_first_200k_trees_match_oracle():
    X, y = datagen.multiclass_2m(200_000)
    cats = datagen.MULTICLASS_CATEGORICAL
    params = dict(objective="multiclass", num_class=10, num_leaves=127, max_bin=255, min_data_in_leaf=20,
                  learning_rate=0.1, categorical_feature=cats)
    ds = lgb.Dataset(X, label=y, params=params)
    ods = O.Dataset(X, y, params)
    assert ds.num_features == ods.num_features
    np.testing.assert_array_equal(ds.bins, ods.bins())
    metas = [ds.feature_meta(f) for f in range(ds.num_features)]
    assert sum(m["bin_type"] == 1 for m in metas) == 8
    assert sum(m["num_bin"] <= 16 for m in metas if m["bin_type"] == 0) >= 12
    b = lgb.Booster(params, ds)
    ob = O.Booster(ods, params)
    equal = _boost_same_gradients(b, ob, 2, K=10, inits=_init_scores(y, "multiclass", 10))
    print(f"[configs] multiclass: {equal} of 20 trees equal", flush=True)
    # categorical splits occur in these trees
    assert any((b.tree(t).decision_type & 1).any() for t in range(20))


@pytest.fixture(scope="module")
def rank_model():
    X, rel, qb = datagen.rank_2m(n_queries=10_000)
    q_train = 2_000
    n_train = int(qb[q_train])
    params = dict(objective="lambdarank", num_leaves=255, max_bin=255, min_data_in_leaf=20, learning_rate=0.1)
    ds = lgb.Dataset(X[:n_train], label=rel[:n_train], group=np.diff(qb[:q_train + 1]), params=params)
    b = lgb.Booster(params, ds)
    print(f"[configs] training 1000 lambdarank trees on {n_train} documents", flush=True)
    for _ in range(1000):
        b.update()
    assert b.num_trees() == 1000
    return X, ds, b


def test_predict_1000_trees_1m_rows_raw_and_binned_equal_oracle(rank_model):
    X, ds, b = rank_model
    assert X.shape[0] >= 1_000_000 and X.shape[1] == 136
    X = X[:1_000_000]
    raw = b.predict(X, raw_score=True)
    bins = lgb.Dataset(X, reference=ds).bins
    binned = b.predict_binned(bins)[:, 0]
    idx = np.arange(0, len(X), 10)
    om = O.LoadedModel(b.model_to_string())
    print(f"[configs] oracle prediction of {len(idx)} rows x 1000 trees", flush=True)
    want = om.predict_raw(X[idx])[:, 0]
    np.testing.assert_array_equal(raw[idx], want)
    np.testing.assert_array_equal(binned[idx], want)
    # every row, not only the sample: binned and raw forms agree exactly
    np.testing.assert_array_equal(binned, raw)


def _leaf_values(b, trees):
    return [b.tree(t).leaf_value for t in trees]


def test_predict_window_and_early_stop_against_a_leaf_loop(rank_model):
    X, ds, b = rank_model
    Xs = X[:20000]
    leaves = b.predict(Xs, pred_leaf=True)            # [rows][1000]
    lv = _leaf_values(b, range(1000))
    om = O.LoadedModel(b.model_to_string())

    def loop(start, num, freq=0, margin=0.0):
        raw = np.zeros(len(Xs))
        active = np.ones(len(Xs), bool)
        for it in range(num):
            t = start + it
            raw = np.where(active, raw + lv[t][leaves[:, t]], raw)
            if freq > 0 and (it + 1) % freq == 0:
                active &= ~(2.0 * np.abs(raw) > margin)
        return raw

    full = loop(0, 1000)
    np.testing.assert_array_equal(b.predict(Xs, raw_score=True), full)
    # a window is the sum over its iterations only; windows add up to the whole
    win = b.predict(Xs, raw_score=True, start_iteration=300, num_iteration=250)
    np.testing.assert_array_equal(win, loop(300, 250))
    np.testing.assert_array_equal(om.predict_window(Xs, 300, 250)[:, 0], win)
    head = b.predict(Xs, raw_score=True, start_iteration=0, num_iteration=300)
    tail = b.predict(Xs, raw_score=True, start_iteration=300, num_iteration=-1)
    np.testing.assert_allclose(head + tail, full, rtol=1e-12, atol=1e-12)
    # early stopping: a row stops once 2|raw| > margin at a multiple of freq iterations
    for freq, margin in ((10, 1.0), (7, 3.0), (50, 0.5)):
        es = b.predict(Xs, raw_score=True, pred_early_stop=True, pred_early_stop_freq=freq,
                       pred_early_stop_margin=margin)
        want = loop(0, 1000, freq, margin)
        np.testing.assert_array_equal(es, want)
        np.testing.assert_array_equal(om.predict_window(Xs, 0, -1, freq, margin)[:, 0], want)
        assert (es != full).any()
