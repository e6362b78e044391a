"""Binning on the device (value_to_bin, binning.rs:637-667) against the host
binning of the same mappers, which the oracle pins: bit-exact bins, and the
trees trained from them identical to the host-binned ones."""
import numpy as np
import pytest

from tests._pkg import DeviceArray, lgb

pytestmark = pytest.mark.gpu


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, 9)).astype(np.float32)
    X[::7, 1] = np.nan                  # NaN bin
    X[::3, 2] = 0.0                     # zero-heavy column
    X[:, 3] = rng.integers(-3, 40, n).astype(np.float32)   # categorical, negatives go to bin 0
    X[:11, 3] = 3e9                     # saturating cast: an unseen category
    X[::5, 3] = np.nan
    X[:, 4] = rng.lognormal(size=n).astype(np.float32)
    X[:, 5] = 1.0                       # trivial column, not used
    X[::2, 6] = np.inf
    X[1::4, 6] = -np.inf
    y = ((X[:, 0] + (X[:, 3] % 5 == 1) + np.nan_to_num(X[:, 1])) > 0.5).astype(np.float32)
    return X, y


@pytest.mark.parametrize("max_bin", [15, 255])
def test_device_bins_equal_host_bins(max_bin):
    X, y = _data(120001, 1)
    params = dict(objective="binary", num_leaves=31, max_bin=max_bin, categorical_feature=[3])
    host = lgb.Dataset(X, label=y, params=params)
    b = lgb.Booster(params, host)
    dx = DeviceArray(X)
    out = DeviceArray(nbytes=len(X) * host.num_features)
    b.bin_rows_device(dx.ptr, len(X), X.shape[1], out.ptr)
    got = out.to_host(np.zeros((len(X), host.num_features), np.uint8))
    np.testing.assert_array_equal(got, host.bins)


@pytest.mark.parametrize("quant", [False, True])
def test_device_binned_training_equals_host_binned(quant):
    X, y = _data(80000, 2)
    params = dict(objective="binary", num_leaves=31, max_bin=63, categorical_feature=[3],
                  use_quantized_grad=quant, num_grad_quant_bins=4)
    host = lgb.Dataset(X, label=y, params=params)
    dev = lgb.Dataset(X, label=y, params=dict(params, device_binning=True))
    bh = lgb.Booster(params, host)
    bd = lgb.Booster(params, dev)
    for _ in range(4):
        bh.update()
        bd.update()
    assert bh.num_trees() == bd.num_trees() == 4
    for t in range(4):
        th, td = bh.tree(t), bd.tree(t)
        for a in ("split_feature", "threshold_in_bin", "threshold", "decision_type", "left_child", "right_child",
                  "leaf_value", "leaf_count"):
            np.testing.assert_array_equal(getattr(th, a), getattr(td, a))
    np.testing.assert_array_equal(bh.train_scores(), bd.train_scores())
    # the stage partition reads the column-major copy built on the device
    rows = np.arange(0, len(X), 3, dtype=np.uint32)
    for b in (bh, bd):
        lib = lgb._capi.lib()
        out = np.zeros(len(rows), np.uint32)
        nl = np.zeros(1, np.int64)
        lgb._capi.check(lib.lgbmhip_partition(b._ctx, lgb.ptr(rows), len(rows), 0, 7, 0, lgb.ptr(out),
                                              lgb.ptr(nl)), b._ctx)
        b._part = (out.copy(), int(nl[0]))
    assert bh._part[1] == bd._part[1]
    np.testing.assert_array_equal(bh._part[0], bd._part[0])
