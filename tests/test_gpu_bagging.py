"""Row bagging (sample_strategy.rs:186-318) on the GPU against the oracle:
the same bag (one generator per 1024-row block, running on from draw to
draw), trees grown on the bag's rows, out-of-bag rows scored by the binned
tree. Tree structure, thresholds and leaf counts exact; leaf values within
FP64 summation noise; scores within 1e-9 relative."""
import numpy as np
import pytest

from tests._pkg import datagen, lgb
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

P = dict(objective="binary", num_leaves=31, learning_rate=0.1, min_data_in_leaf=20, max_bin=63)


def _compare(b, ob, trees):
    for t in range(trees):
        gt, ot = b.tree(t), ob.tree(t)
        assert gt.num_leaves == ot.num_leaves, t
        for a in ("split_feature", "threshold_in_bin", "decision_type", "left_child", "right_child", "leaf_count",
                  "threshold"):
            np.testing.assert_array_equal(getattr(gt, a), getattr(ot, a))
        np.testing.assert_allclose(gt.leaf_value, ot.leaf_value, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("quant,fraction,freq", [(False, 0.3, 2), (False, 0.8, 1), (True, 0.4, 1), (True, 0.7, 3)])
def test_bagging_matches_oracle(quant, fraction, freq):
    X, y = datagen.binary_100k(60000)
    params = dict(P, bagging_fraction=fraction, bagging_freq=freq, bagging_seed=11)
    if quant:
        params.update(use_quantized_grad=True, num_grad_quant_bins=16, quant_train_renew_leaf=True)
    print(f"bagging quant={quant} fraction={fraction} freq={freq}", flush=True)
    b = lgb.Booster(params, lgb.Dataset(X, label=y, params=params))
    ob = O.Booster(O.Dataset(X, y, params), params)
    iters = 6
    for _ in range(iters):
        b.update()
        ob.update()
    bag = ob.bag()
    assert 0 < len(bag) < len(y)
    _compare(b, ob, iters)
    np.testing.assert_allclose(b.train_scores(), ob.scores(), rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("quant", [False, True])
def test_balanced_bagging_matches_oracle(quant):
    X, y = datagen.binary_100k(50000)
    params = dict(P, pos_bagging_fraction=0.3, neg_bagging_fraction=0.6, bagging_freq=1, bagging_seed=5)
    if quant:
        params.update(use_quantized_grad=True, num_grad_quant_bins=16, quant_train_renew_leaf=True)
    b = lgb.Booster(params, lgb.Dataset(X, label=y, params=params))
    ob = O.Booster(O.Dataset(X, y, params), params)
    for _ in range(4):
        b.update()
        ob.update()
    bag = ob.bag()
    pos_rate = (y[bag] > 0).sum() / max((y > 0).sum(), 1)
    assert abs(pos_rate - 0.3) < 0.03
    _compare(b, ob, 4)
    np.testing.assert_allclose(b.train_scores(), ob.scores(), rtol=1e-9, atol=1e-12)


def test_multiclass_bagging_matches_oracle():
    rng = np.random.default_rng(12)
    n, K = 40000, 3
    X = rng.normal(size=(n, 8)).astype(np.float32)
    y = np.argmax(X[:, :K] + rng.gumbel(size=(n, K)), 1).astype(np.float32)
    params = dict(objective="multiclass", num_class=K, num_leaves=15, max_bin=63, bagging_fraction=0.5,
                  bagging_freq=1)
    b = lgb.Booster(params, lgb.Dataset(X, label=y, params=params))
    ob = O.Booster(O.Dataset(X, y, params), params)
    for _ in range(3):
        b.update()
        ob.update()
    _compare(b, ob, 3 * K)
    np.testing.assert_allclose(b.train_scores(), ob.scores(), rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("sampler", ["bagging", "goss"])
def test_quantized_full_data_bag_matches_oracle(sampler):
    # an average bag rate above 0.5 (or GOSS rates above 0.5): the learner
    # keeps every row in view, the quantizer runs over all rows, and the
    # root's sums are quant.rs:144-158's (f32 scales, packed sum by position)
    X, y = datagen.binary_100k(60000)
    params = dict(P, use_quantized_grad=True, num_grad_quant_bins=16, quant_train_renew_leaf=True)
    if sampler == "bagging":
        params.update(bagging_fraction=0.8, bagging_freq=1, bagging_seed=7)
    else:
        params.update(learning_rate=0.5, data_sample_strategy="goss", top_rate=0.4, other_rate=0.3, num_threads=4)
    b = lgb.Booster(params, lgb.Dataset(X, label=y, params=params))
    ob = O.Booster(O.Dataset(X, y, params), params)
    for _ in range(5):
        b.update()
        ob.update()
    bag = ob.bag()
    assert bag is not None and 0.5 * len(y) < len(bag) < len(y)
    _compare(b, ob, 5)
    np.testing.assert_allclose(b.train_scores(), ob.scores(), rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("quant,threads,K", [(False, 4, 1), (True, 1, 1), (False, 3, 3)])
def test_goss_matches_oracle(quant, threads, K):
    rng = np.random.default_rng(14)
    X, y = datagen.binary_100k(40000)
    params = dict(P, learning_rate=0.5, data_sample_strategy="goss", top_rate=0.2, other_rate=0.1,
                  num_threads=threads, bagging_seed=9)
    if K > 1:
        y = np.minimum(np.floor(np.abs(X[:, 0]) * 2), K - 1).astype(np.float32)
        params.update(objective="multiclass", num_class=K)
    if quant:
        params.update(use_quantized_grad=True, num_grad_quant_bins=16, quant_train_renew_leaf=True)
    print(f"goss quant={quant} threads={threads} K={K}", flush=True)
    b = lgb.Booster(params, lgb.Dataset(X, label=y, params=params))
    ob = O.Booster(O.Dataset(X, y, params), params)
    iters = 5   # GOSS starts at iteration 1 / learning_rate = 2
    for _ in range(iters):
        b.update()
        ob.update()
    bag = ob.bag()
    # top 20% by |g h| (more where |g h| ties at the threshold) plus 10% sampled from the rest
    assert bag is not None and 0.25 * len(y) < len(bag) < 0.5 * len(y)
    _compare(b, ob, iters * K)
    np.testing.assert_allclose(b.train_scores(), ob.scores(), rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("goss", [False, True])
def test_lambdarank_row_sampling_matches_oracle(goss):
    # row bagging / GOSS with the ranking objective: gradients of every query,
    # trees on the sampled rows (bagging_by_query is not built)
    X, rel, qb = datagen.rank_2m(400, 300)
    params = dict(objective="lambdarank", num_leaves=31, max_bin=63, learning_rate=0.5)
    if goss:
        params.update(data_sample_strategy="goss", top_rate=0.3, other_rate=0.2, num_threads=2)
    else:
        params.update(bagging_fraction=0.5, bagging_freq=1, bagging_seed=2)
    b = lgb.Booster(params, lgb.Dataset(X, label=rel, group=np.diff(qb), params=params))
    ods = O.Dataset(X, rel, params)
    ods.set_query_boundaries(qb)
    ob = O.Booster(ods, params)
    for _ in range(4):
        b.update()
        ob.update()
    assert ob.bag() is not None
    _compare(b, ob, 4)
    np.testing.assert_allclose(b.train_scores(), ob.scores(), rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("quant", [False, True])
def test_bagging_by_query_matches_oracle(quant):
    # bagging_by_query (sample_strategy.rs:239-276): whole queries in the bag
    X, rel, qb = datagen.rank_2m(600, 300)
    params = dict(objective="lambdarank", num_leaves=31, max_bin=63, learning_rate=0.5, bagging_fraction=0.4,
                  bagging_freq=1, bagging_seed=4, bagging_by_query=True)
    if quant:
        params.update(use_quantized_grad=True, num_grad_quant_bins=16, quant_train_renew_leaf=True)
    b = lgb.Booster(params, lgb.Dataset(X, label=rel, group=np.diff(qb), params=params))
    ods = O.Dataset(X, rel, params)
    ods.set_query_boundaries(qb)
    ob = O.Booster(ods, params)
    for _ in range(4):
        b.update()
        ob.update()
    bag = ob.bag()
    q_of_row = np.searchsorted(qb, bag, side="right") - 1
    sizes = np.diff(qb)
    assert np.array_equal(np.bincount(q_of_row, minlength=len(sizes))[np.unique(q_of_row)],
                          sizes[np.unique(q_of_row)])   # every drawn query whole
    _compare(b, ob, 4)
    np.testing.assert_allclose(b.train_scores(), ob.scores(), rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("bundle", [True, False])
def test_bagging_subset_rule_counts_feature_groups(bundle):
    # 60 dense + 60 30%-non-zero features: 120 features in 61 of the
    # reference's feature groups. With bundling the bag of 40% of the rows
    # takes the subset mode (num_feature_groups < 100, sample_strategy.rs:
    # 198-201): the quantizer draws over the bag's rows; without it (120
    # groups) the full-data mode quantizes every row. Either way the trees are
    # the oracle's, which follows the same rule.
    rng = np.random.default_rng(8)
    n = 30000
    X = np.hstack([rng.normal(size=(n, 60)),
                   rng.normal(size=(n, 60)) * (rng.random((n, 60)) < 0.3)]).astype(np.float32)
    y = (X[:, 0] + X[:, 61] + 0.5 * rng.normal(size=n) > 0).astype(np.float32)
    params = dict(P, bagging_fraction=0.4, bagging_freq=1, bagging_seed=5, use_quantized_grad=True,
                  num_grad_quant_bins=16, quant_train_renew_leaf=True, enable_bundle=bundle)
    print(f"bagging subset rule bundle={bundle}", flush=True)
    ds = lgb.Dataset(X, label=y, params=params)
    od = O.Dataset(X, y, params)
    assert (ds.num_feature_groups, od.num_groups) == ((61, 61) if bundle else (120, 120))
    b = lgb.Booster(params, ds)
    ob = O.Booster(od, params)
    iters = 4
    for _ in range(iters):
        b.update()
        ob.update()
    _compare(b, ob, iters)
    np.testing.assert_allclose(b.train_scores(), ob.scores(), rtol=1e-9, atol=1e-12)
