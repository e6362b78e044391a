"""CPU-side checks of the product: the C ABI library loads and exports every
declared entry point, and host binning equals the oracle's bit for bit."""
import os
import re

import numpy as np
import pytest

from tests._pkg import capi, datagen, lgb
from oracle import pyoracle as O

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "lgbm_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(lgbmhip_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    syms = declared_symbols()
    assert len(syms) >= 30
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(capi.SIGNATURES), set(syms) ^ set(capi.SIGNATURES)


def test_error_convention_without_gpu_call():
    # binning with no rows is InvalidData (dataset.rs:555-592), reported through the ABI
    with pytest.raises(lgb.LightGBMError, match="InvalidData"):
        lgb.Dataset(np.zeros((0, 3), np.float32), label=np.zeros(0, np.float32))
    with pytest.raises(lgb.LightGBMError, match="device_type"):
        lgb.Dataset(np.zeros((3, 3), np.float32), params={"device": "cpu"})


def _compare_bins(X, y, params):
    ds = lgb.Dataset(X, label=y, params=params)
    ods = O.Dataset(X, y, params)
    assert ds.num_features == ods.num_features
    off = 0
    for f in range(ds.num_features):
        m = ds.feature_meta(f)
        o = ods.features[f]
        assert (m["real_index"], m["num_bin"], m["missing_type"], m["default_bin"], m["most_freq_bin"],
                m["bin_type"]) == (o.real, o.num_bin, o.missing_type, o.default_bin, o.most_freq_bin, o.bin_type)
        up = ds.upper[off:off + m["num_bin"]]
        np.testing.assert_array_equal(up, o.upper)
        off += m["num_bin"]
    np.testing.assert_array_equal(ds.bins.astype(np.uint16), ods.bins())
    return ds


@pytest.mark.parametrize("max_bin", [63, 255, 15])
def test_host_binning_matches_oracle(max_bin):
    X, y = datagen.binary_100k(20000)
    _compare_bins(X, y, dict(max_bin=max_bin))


def test_host_binning_missing_zero_skewed_and_sampled():
    rng = np.random.default_rng(3)
    n = 30000
    X = np.stack([rng.normal(size=n), np.exp(rng.normal(size=n)), rng.integers(0, 5, n).astype(float),
                  np.where(rng.random(n) < 0.7, 0.0, rng.normal(size=n)), rng.normal(size=n)], 1).astype(np.float32)
    X[::11, 0] = np.nan
    X[:, 4] = 1.0  # constant -> trivial, dropped
    y = (rng.random(n) > 0.5).astype(np.float32)
    ds = _compare_bins(X, y, dict(max_bin=255, bin_construct_sample_cnt=5000))
    assert ds.num_features == 4


def test_host_binning_categorical_matches_oracle():
    rng = np.random.default_rng(5)
    n = 20000
    z = rng.zipf(1.2, size=n)
    cat = np.where(z <= 50, z - 1, 49).astype(np.float32)
    X = np.stack([rng.normal(size=n).astype(np.float32), cat], 1)
    y = (rng.random(n) > 0.5).astype(np.float32)
    _compare_bins(X, y, dict(max_bin=255, categorical_feature=[1]))


def test_datagen_is_deterministic_and_shaped():
    a = datagen.normals(7, 0, 1000)
    b = datagen.normals(7, 0, 1000)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(datagen.normals(7, 500, 10), a[500:510])
    assert abs(a.mean()) < 0.1 and abs(a.std() - 1) < 0.1
    X, y = datagen.binary_100k(1000)
    assert X.shape == (1000, 28) and set(np.unique(y)) <= {0.0, 1.0}


def test_find_mappers_are_the_bin_dense_mappers():
    rng = np.random.default_rng(40)
    X = rng.normal(size=(30000, 6)).astype(np.float32)
    X[::5, 1] = np.nan
    X[:, 2] = rng.integers(-2, 30, len(X)).astype(np.float32)
    X[:7, 2] = 1e12   # saturating cast: an unseen category
    y = (X[:, 0] > 0).astype(np.float32)
    params = dict(max_bin=63, categorical_feature=[2])
    host = lgb.Dataset(X, label=y, params=params)
    dev = lgb.Dataset(X, label=y, params=dict(params, device_binning=True))
    assert dev._bins is None and dev.num_features == host.num_features
    assert [dev.feature_meta(f) for f in range(dev.num_features)] == \
        [host.feature_meta(f) for f in range(host.num_features)]
    np.testing.assert_array_equal(dev.upper, host.upper)
    np.testing.assert_array_equal(dev.bins, host.bins)   # host apply of the same mappers
    # a mapper-only handle has no bin matrix
    out = np.zeros(dev.bins.size, np.uint8)
    with pytest.raises(lgb.LightGBMError, match="no bin matrix"):
        capi.check(capi.lib().lgbmhip_bins_matrix(dev._handle, capi.ptr(out)))


def test_feature_infos_are_sample_min_max_and_categories():
    rng = np.random.default_rng(41)
    X = rng.normal(size=(2000, 4)).astype(np.float32)
    X[:, 1] = np.abs(X[:, 1]) + 1.0       # positive only: 0 is not a sample value (no zeros)
    X[:, 2] = 7.0                          # constant: unused
    X[:, 3] = rng.integers(0, 4, len(X)).astype(np.float32)
    ds = lgb.Dataset(X, label=(X[:, 0] > 0).astype(np.float32), params=dict(categorical_feature=[3]))
    infos = ds.feature_infos().split(" ")
    assert infos[0] == "[%.17g:%.17g]" % (float(X[:, 0].min()), float(X[:, 0].max()))
    assert infos[1] == "[%.17g:%.17g]" % (float(X[:, 1].min()), float(X[:, 1].max()))
    assert infos[2] == "none"
    # the categories of the bins in bin order, bin 0 holding -1 (binning.rs:704-779 order [-1, 3, 1, 0, 2])
    cats = [int(c) for c in infos[3].split(":")]
    assert cats[0] == -1 and sorted(cats[1:]) == [0, 1, 2, 3]


def _groups(X, **params):
    params = dict(dict(objective="binary"), **params)
    ds = lgb.Dataset(X, label=np.zeros(len(X), np.float32), params=params)
    od = O.Dataset(X, np.zeros(len(X), np.float32), params)
    return ds.num_features, ds.num_feature_groups, od.num_groups


def test_feature_group_count_cases_the_greedy_bundling_fixes():
    # feature_groups.rs:62-158: a feature joins a group only when its sampled
    # non-zero positions conflict in at most total/10000 of them; groups used
    # by under 40% of the sample merge into one trailing group.
    rng = np.random.default_rng(5)
    n = 30000
    dense = rng.normal(size=(n, 12)).astype(np.float32)
    # dense columns conflict everywhere: one group each
    assert _groups(dense) == (12, 12, 12)
    # mutually exclusive one-hot columns share one group (it covers every row)
    onehot = np.zeros((n, 9), np.float32)
    onehot[np.arange(n), rng.integers(0, 9, n)] = 1.0
    assert _groups(onehot) == (9, 1, 1)
    assert _groups(np.hstack([dense, onehot])) == (21, 13, 13)
    # independent 30%-non-zero columns conflict in ~9% of rows: each its own
    # group in the first pass, all under 40% used, merged into one
    sparse = (rng.normal(size=(n, 10)) * (rng.random((n, 10)) < 0.3)).astype(np.float32)
    assert _groups(np.hstack([dense, sparse])) == (22, 13, 13)
    # without bundling every feature is a group
    assert _groups(np.hstack([dense, sparse]), enable_bundle=False) == (22, 22, 22)


def test_feature_group_count_decides_the_bagging_subset_rule():
    # 60 dense + 60 sparse features: 120 features but 61 groups, so a bag of
    # at most half the rows takes the reference's subset mode (num_feature_groups
    # < 100, sample_strategy.rs:198-201); the count reaches the engine unchanged
    rng = np.random.default_rng(6)
    n = 20000
    X = np.hstack([rng.normal(size=(n, 60)),
                   rng.normal(size=(n, 60)) * (rng.random((n, 60)) < 0.3)]).astype(np.float32)
    nf, ng, og = _groups(X)
    assert (nf, ng, og) == (120, 61, 61)
