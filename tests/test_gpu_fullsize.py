"""Properties at the headline size, 10.5M x 28 (binary, num_leaves 255,
max_bin 255, int8-quantized gradients with 16 levels), where the oracle is
too slow to train beside the engine:

  * every tree's leaf counts add up to the rows;
  * the training scores equal the trained model's raw prediction of the
    training rows (the score update and prediction walk the same trees);
  * two runs give bit-identical trees and scores (determinism);
  * the device binning pass gives the host binning's bins;
  * the GPU AUC of the training scores equals the oracle's exactly.
"""
import numpy as np
import pytest

from tests._pkg import DeviceArray, datagen, lgb
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

N = 10_500_000


@pytest.fixture(scope="module")
def data():
    X, y = datagen.binary_10p5m(N)
    return X, y


def _train(ds, iters):
    params = dict(objective="binary", num_leaves=255, max_bin=255, min_data_in_leaf=20, learning_rate=0.1,
                  use_quantized_grad=True, num_grad_quant_bins=16, stochastic_rounding=True,
                  quant_train_renew_leaf=True, metric="auc")
    b = lgb.Booster(params, ds)
    for _ in range(iters):
        b.update()
    return b


def test_headline_size_properties(data):
    X, y = data
    print("binning 10.5M x 28", flush=True)
    ds = lgb.Dataset(X, label=y, params=dict(max_bin=255, device_binning=True))
    b = _train(ds, 3)
    trees = [b.tree(t) for t in range(b.num_trees())]
    assert len(trees) == 3 and all(t.num_leaves == 255 for t in trees)
    for t in trees:
        assert int(t.leaf_count.sum()) == N
    scores = b.train_scores()
    rows = np.random.default_rng(0).choice(N, 200_000, replace=False)
    np.testing.assert_allclose(b.predict(X[rows], raw_score=True), scores[rows], rtol=0, atol=1e-12)
    # determinism
    b2 = _train(ds, 3)
    for t in range(3):
        for a in ("split_feature", "threshold_in_bin", "leaf_value", "leaf_count"):
            np.testing.assert_array_equal(getattr(b.tree(t), a), getattr(b2.tree(t), a))
    np.testing.assert_array_equal(scores, b2.train_scores())
    # GPU AUC of the training scores against the oracle's (integer sums: exact)
    auc = b.eval_train()[0].metric_value
    assert auc == O.metric_auc(scores, y)


def test_headline_size_device_bins(data):
    X, y = data
    ds = lgb.Dataset(X, label=y, params=dict(max_bin=255, device_binning=True))
    b = lgb.Booster(dict(objective="binary"), ds)
    dx = DeviceArray(X)
    out = DeviceArray(nbytes=N * ds.num_features)
    b.bin_rows_device(dx.ptr, N, X.shape[1], out.ptr)
    got = out.to_host(np.zeros((N, ds.num_features), np.uint8))
    np.testing.assert_array_equal(got, ds.bins)   # host binning with the same mappers
