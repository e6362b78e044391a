"""Parity of the HIP engine (through the C ABI) with the CPU oracle.

Tolerances:
  * integer / index work (partition, quantized histograms, discretization,
    split thresholds on identical histograms, prediction): bit-exact;
  * FP64 histograms: the engine sums exact 64-bit fixed-point values, the
    oracle sums sequentially in FP64; each bin may differ by at most
    (count_bin + 1) * 2^-(s+1) + 64 ulp(|bin|), with 2^-s the fixed-point grid;
  * gradients: libm exp vs the device exp may differ in the last ulp of the
    f64 response, so at most 1e-5 of rows may differ, each by <= 1 f32 ulp.
"""
import numpy as np
import pytest

from tests._pkg import datagen, lgb, small_binary
from tests._trees import compare_or_near_tie
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

P_BIN = dict(objective="binary", num_leaves=31, learning_rate=0.1, min_data_in_leaf=20, max_bin=63)


def _booster(X, y, params):
    ds = lgb.Dataset(X, label=y, params=params)
    return ds, lgb.Booster(params, ds)


def _fix_shift(count, m):
    r = 1
    while (1 << r) <= count and r < 62:
        r += 1
    e = 0
    if m > 0:
        e = int(np.frexp(np.float32(m))[1])
    return 62 - r - e


def test_histogram_f64_within_fixed_point_bound():
    X, y = datagen.binary_100k(50000)
    ds, b = _booster(X, y, P_BIN)
    rng = np.random.default_rng(1)
    g = rng.normal(size=len(y)).astype(np.float32)
    h = rng.random(len(y)).astype(np.float32)
    rows = np.sort(rng.choice(len(y), 20000, replace=False)).astype(np.uint32)
    B = max(ds.feature_meta(f)["num_bin"] for f in range(ds.num_features))
    out = np.zeros(ds.num_features * B * 2, np.float64)
    nb = np.zeros(1, np.int32)
    lib = lgb._capi.lib()
    lgb._capi.check(lib.lgbmhip_build_histogram(b._ctx, lgb.ptr(rows), len(rows), lgb.ptr(g), lgb.ptr(h),
                                                 lgb.ptr(out), lgb.ptr(nb)), b._ctx)
    assert nb[0] == B
    gpu = out.reshape(ds.num_features, B, 2)
    mfb = np.array([ds.feature_meta(f)["most_freq_bin"] for f in range(ds.num_features)], np.int32)
    cpu = O.hist_f64(ds.bins, mfb, B, rows, g, h)
    sg = _fix_shift(len(rows), np.abs(g[rows]).max())
    sh = _fix_shift(len(rows), np.abs(h[rows]).max())
    cnt = np.zeros((ds.num_features, B))
    for f in range(ds.num_features):
        cnt[f] = np.bincount(ds.bins[rows, f], minlength=B)
    for c, s in ((0, sg), (1, sh)):
        bound = (cnt + 1) * 2.0 ** -(s + 1) + 64 * np.spacing(np.abs(cpu[:, :, c]))
        assert np.all(np.abs(gpu[:, :, c] - cpu[:, :, c]) <= bound)
    # the most frequent bins are never accumulated
    for f in range(ds.num_features):
        assert gpu[f, mfb[f]].tolist() == [0.0, 0.0]


def test_histogram_is_deterministic_and_order_free():
    X, y = datagen.binary_100k(30000)
    ds, b = _booster(X, y, P_BIN)
    rng = np.random.default_rng(2)
    g = rng.normal(size=len(y)).astype(np.float32)
    h = rng.random(len(y)).astype(np.float32)
    rows = np.arange(len(y), dtype=np.uint32)
    B = max(ds.feature_meta(f)["num_bin"] for f in range(ds.num_features))
    lib = lgb._capi.lib()
    outs = []
    for perm in (rows, rows[::-1].copy(), rng.permutation(rows).astype(np.uint32)):
        out = np.zeros(ds.num_features * B * 2, np.float64)
        nb = np.zeros(1, np.int32)
        lgb._capi.check(lib.lgbmhip_build_histogram(b._ctx, lgb.ptr(perm), len(perm), lgb.ptr(g), lgb.ptr(h),
                                                     lgb.ptr(out), lgb.ptr(nb)), b._ctx)
        outs.append(out)
    assert all(np.array_equal(outs[0], o) for o in outs[1:])


# (max |gq|, max hq): the 16-bin quantizer's range (32-bit LDS bins, folded
# mid-chunk), a range that folds after every pass, and full int8 values (a
# negative hq is a u8 up to 255: the 64-bit LDS bins)
@pytest.mark.parametrize("gmax,hmax", [(8, 16), (16, 31), (128, 127)])
def test_histogram_int_bit_exact(gmax, hmax):
    X, y = datagen.binary_100k(40000)
    params = dict(P_BIN, use_quantized_grad=True, num_grad_quant_bins=16)
    ds, b = _booster(X, y, params)
    rng = np.random.default_rng(3)
    gq = rng.integers(-gmax, min(gmax, 127) + 1, len(y)).astype(np.int8)
    hq = rng.integers(-128 if gmax == 128 else 0, hmax + 1, len(y)).astype(np.int8)
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


def _random_hist(rng, nb, mfb, scale=1.0):
    cnt = rng.integers(0, 60, nb).astype(np.float64)
    g = rng.normal(size=nb) * cnt * scale
    h = cnt * (0.05 + 0.2 * rng.random(nb))
    hist = np.stack([g, h], 1)
    tot_g, tot_h = g.sum(), h.sum()
    hist[mfb] = 0.0
    return hist, tot_g, tot_h, int(cnt.sum())


@pytest.mark.parametrize("missing_type", [0, 2])
def test_split_search_bit_exact_on_identical_histograms(missing_type):
    X, y = small_binary(4000, 6, nan_col=0 if missing_type == 2 else None)
    params = dict(P_BIN, min_data_in_leaf=5, max_bin=255)
    ds, b = _booster(X, y, params)
    rng = np.random.default_rng(4)
    lib = lgb._capi.lib()
    for f in range(ds.num_features):
        m = ds.feature_meta(f)
        if m["missing_type"] != missing_type:
            continue
        for trial in range(20):
            hist, sg, sh, nd = _random_hist(rng, m["num_bin"], m["most_freq_bin"])
            od = np.zeros(7)
            oi = np.zeros(5, np.int32)
            lgb._capi.check(lib.lgbmhip_find_best_split(b._ctx, f, lgb.ptr(np.ascontiguousarray(hist)), sg, sh, nd,
                                                         0.0, lgb.ptr(od), lgb.ptr(oi), None), b._ctx)
            r = O.find_best_threshold(hist, m["num_bin"], m["missing_type"], m["default_bin"], m["most_freq_bin"],
                                      m["bin_type"], params, sg, sh, nd)
            assert bool(oi[4]) == r["splittable"]
            if r["gain"] == -np.inf:
                assert od[0] == -np.inf
                continue
            assert od[0] == r["gain"] and oi[0] == r["threshold"] and bool(oi[1]) == r["default_left"]
            assert (oi[2], oi[3]) == (r["left_count"], r["right_count"])
            assert od[1] == r["left_output"] and od[2] == r["right_output"]
            assert (od[3], od[4], od[5], od[6]) == (r["left_sum_gradient"], r["left_sum_hessian"],
                                                    r["right_sum_gradient"], r["right_sum_hessian"])


def test_split_search_int_bit_exact():
    X, y = small_binary(4000, 6)
    params = dict(P_BIN, min_data_in_leaf=5, max_bin=63, use_quantized_grad=True, num_grad_quant_bins=16)
    ds, b = _booster(X, y, params)
    rng = np.random.default_rng(5)
    lib = lgb._capi.lib()
    for f in range(ds.num_features):
        m = ds.feature_meta(f)
        for trial in range(20):
            nb = m["num_bin"]
            cnt = rng.integers(0, 50, nb)
            gs = rng.integers(-8, 9, nb) * cnt // 3
            hs = cnt * 7
            packed = (gs.astype(np.int64) << 32) + hs.astype(np.int64)
            tot = int(packed.sum())
            packed[m["most_freq_bin"]] = 0
            gscale, hscale = 0.01, 0.002
            od = np.zeros(7)
            oi = np.zeros(5, np.int32)
            og = np.zeros(2, np.int64)
            lgb._capi.check(lib.lgbmhip_find_best_split_int(b._ctx, f, lgb.ptr(packed), tot, gscale, hscale,
                                                             int(cnt.sum()), 0.0, lgb.ptr(od), lgb.ptr(oi),
                                                             lgb.ptr(og), None), b._ctx)
            r = O.find_best_threshold_int(packed, nb, m["missing_type"], m["default_bin"], m["most_freq_bin"],
                                          params, tot, gscale, hscale, int(cnt.sum()))
            assert bool(oi[4]) == r["splittable"]
            if r["gain"] == -np.inf:
                continue
            assert od[0] == r["gain"] and oi[0] == r["threshold"]
            assert (oi[2], oi[3]) == (r["left_count"], r["right_count"])
            assert (int(og[0]), int(og[1])) == (r["left_sum_gh"], r["right_sum_gh"])


def _bits_to_bins(bits):
    return sorted(w * 32 + k for w in range(len(bits)) for k in range(32) if (int(bits[w]) >> k) & 1)


def _categorical_stage_booster(quant):
    """Three categorical columns (3, 40 and 12 categories: one-hot and sorted
    searches) and one numerical column."""
    rng = np.random.default_rng(7)
    n = 6000
    X = np.column_stack([rng.integers(0, 3, n), rng.integers(0, 40, n), rng.integers(0, 12, n),
                         rng.normal(size=n)]).astype(np.float64)
    y = (rng.random(n) < 0.5).astype(np.float64)
    params = dict(P_BIN, min_data_in_leaf=5, max_bin=255, categorical_feature=[0, 1, 2], min_data_per_group=10,
                  cat_smooth=5.0)
    if quant:
        params.update(use_quantized_grad=True, num_grad_quant_bins=16)
    ds, b = _booster(X, y, params)
    return ds, b, params


def test_categorical_split_search_bit_exact():
    # split.rs:446-670 on identical FP64 histograms
    ds, b, params = _categorical_stage_booster(False)
    rng = np.random.default_rng(8)
    lib = lgb._capi.lib()
    seen = 0
    for f in range(ds.num_features):
        m = ds.feature_meta(f)
        if m["bin_type"] != 1:
            continue
        for trial in range(20):
            hist, sg, sh, nd = _random_hist(rng, m["num_bin"], m["most_freq_bin"])
            od = np.zeros(7)
            oi = np.zeros(5, np.int32)
            cb = np.zeros(8, np.uint32)
            lgb._capi.check(lib.lgbmhip_find_best_split(b._ctx, f, lgb.ptr(np.ascontiguousarray(hist)), sg, sh, nd,
                                                         0.0, lgb.ptr(od), lgb.ptr(oi), lgb.ptr(cb)), b._ctx)
            r = O.find_best_threshold(hist, m["num_bin"], m["missing_type"], m["default_bin"], m["most_freq_bin"],
                                      1, params, sg, sh, nd)
            assert bool(oi[4]) == r["splittable"]
            if not r["splittable"]:
                continue
            seen += 1
            assert od[0] == r["gain"] and not oi[1]
            assert (oi[2], oi[3]) == (r["left_count"], r["right_count"])
            assert od[1] == r["left_output"] and od[2] == r["right_output"]
            assert _bits_to_bins(cb) == sorted(r["cat_threshold"])
    assert seen >= 20


def test_categorical_split_search_int_bit_exact():
    # split.rs:1194-1400 on identical packed histograms
    ds, b, params = _categorical_stage_booster(True)
    rng = np.random.default_rng(9)
    lib = lgb._capi.lib()
    seen = 0
    for f in range(ds.num_features):
        m = ds.feature_meta(f)
        if m["bin_type"] != 1:
            continue
        for trial in range(20):
            nb = m["num_bin"]
            cnt = rng.integers(0, 50, nb)
            gs = rng.integers(-8, 9, nb) * cnt // 3
            hs = cnt * 7
            packed = (gs.astype(np.int64) << 32) + hs.astype(np.int64)
            tot = int(packed.sum())
            packed[m["most_freq_bin"]] = 0
            gscale, hscale = 0.01, 0.002
            od = np.zeros(7)
            oi = np.zeros(5, np.int32)
            og = np.zeros(2, np.int64)
            cb = np.zeros(8, np.uint32)
            lgb._capi.check(lib.lgbmhip_find_best_split_int(b._ctx, f, lgb.ptr(packed), tot, gscale, hscale,
                                                             int(cnt.sum()), 0.0, lgb.ptr(od), lgb.ptr(oi),
                                                             lgb.ptr(og), lgb.ptr(cb)), b._ctx)
            r = O.find_best_threshold_int(packed, nb, m["missing_type"], m["default_bin"], m["most_freq_bin"],
                                          params, tot, gscale, hscale, int(cnt.sum()), bin_type=1)
            assert bool(oi[4]) == r["splittable"]
            if not r["splittable"]:
                continue
            seen += 1
            assert od[0] == r["gain"] and not oi[1]
            assert (oi[2], oi[3]) == (r["left_count"], r["right_count"])
            assert od[1] == r["left_output"] and od[2] == r["right_output"]
            assert (int(og[0]), int(og[1])) == (r["left_sum_gh"], r["right_sum_gh"])
            assert _bits_to_bins(cb) == sorted(r["cat_threshold"])
    assert seen >= 20


def test_partition_is_stable_and_exact():
    X, y = small_binary(20000, 5, nan_col=2)
    ds, b = _booster(X, y, dict(P_BIN, max_bin=255))
    rng = np.random.default_rng(6)
    lib = lgb._capi.lib()
    for f in range(ds.num_features):
        m = ds.feature_meta(f)
        rows = np.sort(rng.choice(len(y), 12345, replace=False)).astype(np.uint32)
        thr = int(rng.integers(0, m["num_bin"]))
        for dl in (0, 1):
            out = np.zeros_like(rows)
            nl = np.zeros(1, np.int64)
            lgb._capi.check(lib.lgbmhip_partition(b._ctx, lgb.ptr(rows), len(rows), f, thr, dl, lgb.ptr(out),
                                                   lgb.ptr(nl)), b._ctx)
            bins = ds.bins[rows, f].astype(int)
            miss = (m["missing_type"] == 2) & (bins == m["num_bin"] - 1)
            left = np.where(miss, bool(dl), bins <= thr)
            np.testing.assert_array_equal(out, np.concatenate([rows[left], rows[~left]]))
            assert nl[0] == left.sum()


def test_add_leaf_outputs_follows_the_partition():
    # partition.rs:178-191: score[i] += leaf_values[leaf of row i], with the
    # rows' leaves those of the oracle's tree grown on the same gradients
    X, y = datagen.binary_100k(50000)
    params = dict(P_BIN, num_leaves=63, max_bin=255)
    ds, b = _booster(X, y, params)
    ob = O.Booster(O.Dataset(X, y, params), params)
    rng = np.random.default_rng(16)
    g, h = O.binary_gradients(rng.normal(size=len(y)), y)
    b.train_tree(g, h)
    ob.train_tree(g, h)
    nl = b.tree(0).num_leaves
    lv = rng.normal(size=nl)
    score = rng.normal(size=len(y))
    out = score.copy()
    lib = lgb._capi.lib()
    lgb._capi.check(lib.lgbmhip_add_leaf_outputs(b._ctx, lgb.ptr(lv), nl, lgb.ptr(out)), b._ctx)
    np.testing.assert_array_equal(out, score + lv[ob.leaf_of_row()])
    st = lib.lgbmhip_add_leaf_outputs(b._ctx, lgb.ptr(lv), nl - 1, lgb.ptr(out))
    assert st == 1


def test_binary_gradients_match_oracle():
    X, y = datagen.binary_100k(200000)
    ds, b = _booster(X, y, P_BIN)
    rng = np.random.default_rng(7)
    s = rng.normal(size=len(y)) * 3
    g = np.zeros(len(y), np.float32)
    h = np.zeros(len(y), np.float32)
    lib = lgb._capi.lib()
    lgb._capi.check(lib.lgbmhip_gradients(b._ctx, lgb.ptr(s), lgb.ptr(g), lgb.ptr(h)), b._ctx)
    og, oh = O.binary_gradients(s, y)
    for a, c in ((g, og), (h, oh)):
        diff = a != c
        assert diff.sum() <= max(1, int(1e-5 * len(y)))
        if diff.any():
            assert np.all(np.abs(a[diff].view(np.int32) - c[diff].view(np.int32)) <= 1)


def test_multiclass_gradients_match_oracle():
    rng = np.random.default_rng(8)
    n, K = 50000, 5
    X = rng.normal(size=(n, 6)).astype(np.float32)
    y = rng.integers(0, K, n).astype(np.float32)
    ds, b = _booster(X, y, dict(P_BIN, objective="multiclass", num_class=K))
    s = rng.normal(size=n * K)
    g = np.zeros(n * K, np.float32)
    h = np.zeros(n * K, np.float32)
    lib = lgb._capi.lib()
    lgb._capi.check(lib.lgbmhip_gradients(b._ctx, lgb.ptr(s), lgb.ptr(g), lgb.ptr(h)), b._ctx)
    og, oh = O.multiclass_gradients(s, y, K)
    for a, c in ((g, og), (h, oh)):
        diff = a != c
        assert diff.sum() <= max(2, int(1e-4 * len(a)))
        if diff.any():
            assert np.all(np.abs(a[diff].view(np.int32) - c[diff].view(np.int32)) <= 1)


def test_discretize_bit_exact():
    X, y = datagen.binary_100k(30000)
    params = dict(P_BIN, use_quantized_grad=True, num_grad_quant_bins=16, stochastic_rounding=True)
    ds, b = _booster(X, y, params)
    rng = np.random.default_rng(9)
    g = rng.normal(size=len(y)).astype(np.float32) * 0.3
    h = (rng.random(len(y)) * 0.25).astype(np.float32)
    lib = lgb._capi.lib()
    gq = np.zeros(len(y), np.int8)
    hq = np.zeros(len(y), np.int8)
    sc = np.zeros(2)
    lgb._capi.check(lib.lgbmhip_discretize(b._ctx, lgb.ptr(g), lgb.ptr(h), 3, lgb.ptr(gq), lgb.ptr(hq),
                                            lgb.ptr(sc)), b._ctx)
    ogq, ohq, gs, hs = O.discretize(g, h, 16, seed=0, stochastic=True, num_calls=3)
    assert (sc[0], sc[1]) == (gs, hs)
    np.testing.assert_array_equal(gq, ogq)
    np.testing.assert_array_equal(hq, ohq)


def _compare_tree(gt, ot, value_rtol=1e-9):
    """Structure must match exactly; values within FP64-summation noise."""
    assert gt.num_leaves == ot.num_leaves
    np.testing.assert_array_equal(gt.split_feature, ot.split_feature)
    np.testing.assert_array_equal(gt.threshold_in_bin, ot.threshold_in_bin)
    np.testing.assert_array_equal(gt.decision_type, ot.decision_type)
    np.testing.assert_array_equal(gt.left_child, ot.left_child)
    np.testing.assert_array_equal(gt.right_child, ot.right_child)
    np.testing.assert_array_equal(gt.leaf_count, ot.leaf_count)
    np.testing.assert_array_equal(gt.threshold, ot.threshold)
    np.testing.assert_allclose(gt.leaf_value, ot.leaf_value, rtol=value_rtol, atol=1e-12)


def test_learner_tree_matches_oracle_on_same_gradients():
    X, y = datagen.binary_100k(100000)
    params = dict(P_BIN, num_leaves=31, max_bin=255)
    ds, b = _booster(X, y, params)
    ob = O.Booster(O.Dataset(X, y, params), params)
    rng = np.random.default_rng(10)
    trees_equal = 0
    for t in range(4):
        s = rng.normal(size=len(y)) * 0.5
        g, h = O.binary_gradients(s, y)
        b.train_tree(g, h)
        ob.train_tree(g, h)
        _compare_tree(b.tree(t), ob.tree(t))
        trees_equal += 1
        np.testing.assert_array_equal(b.leaf_of_row(), ob.leaf_of_row())
    assert trees_equal == 4


def test_quantized_learner_tree_bit_exact_structure():
    X, y = datagen.binary_100k(60000)
    params = dict(P_BIN, num_leaves=31, max_bin=63, use_quantized_grad=True, num_grad_quant_bins=16,
                  quant_train_renew_leaf=True)
    ds, b = _booster(X, y, params)
    ob = O.Booster(O.Dataset(X, y, params), params)
    rng = np.random.default_rng(11)
    for t in range(3):
        s = rng.normal(size=len(y)) * 0.5
        g, h = O.binary_gradients(s, y)
        b.train_tree(g, h)
        ob.train_tree(g, h)
        # integer histograms are exact, so the structure and the split sums are too
        _compare_tree(b.tree(t), ob.tree(t), value_rtol=1e-9)


def _same_trees_as_oracle(params, n=100000, trees=2, seed=20, tuning=None, value_rtol=1e-9):
    """Trains `trees` trees on the same gradients on both sides; every tree
    must equal the oracle's and every row must land in the same leaf."""
    X, y = datagen.binary_100k(n)
    ds, b = _booster(X, y, params)
    if tuning:
        b.set_tuning(**tuning)
    op = {k: v for k, v in params.items() if k != "use_graph"}   # an engine option, not a model parameter
    ob = O.Booster(O.Dataset(X, y, op), op)
    rng = np.random.default_rng(seed)
    for t in range(trees):
        s = rng.normal(size=len(y)) * 0.5
        g, h = O.binary_gradients(s, y)
        b.train_tree(g, h)
        ob.train_tree(g, h)
        _compare_tree(b.tree(t), ob.tree(t), value_rtol=value_rtol)
        np.testing.assert_array_equal(b.leaf_of_row(), ob.leaf_of_row())
    return b


@pytest.mark.parametrize("quant", [False, True])
def test_batched_growth_255_leaves_matches_oracle(quant):
    # many rounds, up to 32 splits each, splits made ahead of the sequential
    # order and a tree that fills up with splits still pending
    params = dict(P_BIN, num_leaves=255, max_bin=63, min_data_in_leaf=20)
    if quant:
        params.update(use_quantized_grad=True, num_grad_quant_bins=16, quant_train_renew_leaf=True)
    b = _same_trees_as_oracle(params)
    assert b.tree(0).num_leaves == 255


def test_growth_rounds_per_launch_do_not_change_trees():
    # one round per launch: every round after the first runs as a continuation
    params = dict(P_BIN, num_leaves=127, max_bin=63)
    _same_trees_as_oracle(params, trees=2, tuning=dict(rounds_main=1, rounds_cont=1))
    # stream launches, uneven batches of rounds
    _same_trees_as_oracle(dict(params, use_graph=False), trees=1, tuning=dict(rounds_main=2, rounds_cont=3))


# Launch shapes of the quantized histogram: one 28-feature group at 8 lanes
# per row (a 64 KiB LDS budget), leaves cut into few long chunks that fold
# their 32-bit bins into int64 sums several times (8 chunks per leaf), and
# short chunks (1024 rows). Trees must not depend on any of them.
@pytest.mark.parametrize("tuning", [dict(hist_lds_budget=65536), dict(hist_chunks=8), dict(hist_min_rows=1024),
                                    dict(late_round=1), dict(late_round=0)])
def test_quantized_histogram_launch_shapes_do_not_change_trees(tuning):
    params = dict(P_BIN, num_leaves=63, max_bin=255, min_data_in_leaf=20, use_quantized_grad=True,
                  num_grad_quant_bins=16, quant_train_renew_leaf=True)
    b = _same_trees_as_oracle(params, trees=2, tuning=tuning)
    if "late_round" in tuning:
        assert b.get_tuning()["late_round"] == tuning["late_round"]


# late_round 1: every growth round after the first searches straight from
# the histogram pass's row-chunk units (FP64 pairs here, many chunks per
# leaf early in the tree), with no slab reduction
def test_fp64_search_from_histogram_units_matches_oracle():
    params = dict(P_BIN, num_leaves=127, max_bin=63, min_data_in_leaf=20)
    _same_trees_as_oracle(params, trees=2, tuning=dict(late_round=1, hist_chunks=16))


@pytest.mark.parametrize("quant", [False, True])
def test_search_from_histogram_units_same_boosted_model(quant):
    """Whole runs with every round, no round and the default rounds searching
    from the histogram units give one model bit for bit."""
    X, y = datagen.binary_100k(120000)
    params = dict(P_BIN, num_leaves=255, max_bin=255, min_data_in_leaf=10)
    if quant:
        params.update(use_quantized_grad=True, num_grad_quant_bins=16, stochastic_rounding=True,
                      quant_train_renew_leaf=True)
    models = []
    for late in (0, 1, -1):
        ds, b = _booster(X, y, params)
        b.set_tuning(late_round=late)
        for _ in range(12):
            b.update()
        models.append(b)
    assert models[2].get_tuning()["late_round"] == 12
    for m in models[1:]:
        assert m.model_to_string() == models[0].model_to_string()
        np.testing.assert_array_equal(m.train_scores(), models[0].train_scores())
    for m in models:
        m.free()


def test_batched_growth_with_max_depth_and_min_data():
    # children that may not be searched (depth, min_data) inside batched rounds
    params = dict(P_BIN, num_leaves=255, max_bin=63, max_depth=6, min_data_in_leaf=300)
    b = _same_trees_as_oracle(params, trees=2)
    assert b.tree(0).num_leaves < 255


def test_quantized_boosting_255_leaves_matches_oracle():
    X, y = datagen.binary_100k(100000)
    params = dict(P_BIN, num_leaves=255, max_bin=63, use_quantized_grad=True, num_grad_quant_bins=16,
                  stochastic_rounding=True, quant_train_renew_leaf=True)
    ds, b = _booster(X, y, params)
    ob = O.Booster(O.Dataset(X, y, params), params)
    for _ in range(3):
        b.update()
        ob.update()
    for t in range(3):
        _compare_tree(b.tree(t), ob.tree(t), value_rtol=1e-9)
    np.testing.assert_allclose(b.train_scores(), ob.scores(), rtol=1e-9, atol=1e-12)


def _boost_both(b, ob, n_iter, K=1):
    """n_iter iterations on both sides, each computing its own gradients.
    Every tree must equal the oracle's; a differing split is accepted only as
    a reported near-tie (equal f32 gains), after which the runs diverge and
    the comparison stops. Returns the iterations compared."""
    for it in range(n_iter):
        b.update()
        ob.update()
        for k in range(K):
            if not compare_or_near_tie(b.tree(it * K + k), ob.tree(it * K + k)):
                return it
    return n_iter


def test_boosting_matches_oracle_and_predicts():
    X, y = datagen.binary_100k(100000)
    params = dict(P_BIN, num_leaves=31, max_bin=255)
    ds, b = _booster(X, y, params)
    ob = O.Booster(O.Dataset(X, y, params), params)
    n_iter = 10
    done = _boost_both(b, ob, n_iter)
    s_gpu, s_cpu = b.train_scores(), ob.scores()
    if done == n_iter:
        np.testing.assert_allclose(s_gpu, s_cpu, rtol=1e-9, atol=1e-12)
    raw = b.predict(X, raw_score=True)
    # training scores equal raw predictions (end_to_end.rs:53-57)
    np.testing.assert_array_equal(raw, s_gpu)
    # binned prediction equals raw prediction on training rows
    np.testing.assert_array_equal(b.predict_binned(ds.bins)[:, 0], raw)
    assert abs(O.auc(s_gpu, y) - O.auc(s_cpu, y)) < 1e-4


def test_training_is_deterministic():
    X, y = datagen.binary_100k(50000)
    params = dict(P_BIN, num_leaves=63, max_bin=255)
    runs = []
    for _ in range(2):
        ds, b = _booster(X, y, params)
        for _ in range(5):
            b.update()
        runs.append(b.train_scores())
        b.free()
    np.testing.assert_array_equal(runs[0], runs[1])


def test_graph_and_stream_paths_agree():
    X, y = datagen.binary_100k(40000)
    out = []
    for use_graph in (False, True):
        params = dict(P_BIN, num_leaves=31, use_graph=use_graph)
        ds, b = _booster(X, y, params)
        for _ in range(4):
            b.update()
        out.append(b.train_scores())
        b.free()
    np.testing.assert_array_equal(out[0], out[1])


def test_quantized_boosting_quality_close_to_plain():
    X, y = datagen.binary_100k(100000)
    Xv, yv = datagen.holdout("binary_100k", 20000)
    aucs = {}
    for q in (False, True):
        params = dict(P_BIN, num_leaves=31, max_bin=63, use_quantized_grad=q, num_grad_quant_bins=16,
                      quant_train_renew_leaf=True)
        ds, b = _booster(X, y, params)
        # past the first steep iterations (at 20 the AUC still climbs 0.89 -> 0.93 and
        # one iteration's noise moves it by ~1e-3)
        for _ in range(60):
            b.update()
        aucs[q] = O.auc(b.predict(Xv, raw_score=True), yv)
    assert abs(aucs[True] - aucs[False]) <= 1e-3, aucs


def test_multiclass_boosting_trains():
    rng = np.random.default_rng(12)
    n, K = 30000, 4
    X = rng.normal(size=(n, 8)).astype(np.float32)
    W = rng.normal(size=(8, K))
    y = np.argmax(X @ W + rng.gumbel(size=(n, K)), 1).astype(np.float32)
    params = dict(objective="multiclass", num_class=K, num_leaves=15, max_bin=63)
    ds, b = _booster(X, y, params)
    ob = O.Booster(O.Dataset(X, y, params), params)
    for it in range(3):
        b.update()
        ob.update()
    for t in range(3 * K):
        if not compare_or_near_tie(b.tree(t), ob.tree(t)):
            break
    p = b.predict(X)
    assert p.shape == (n, K) and np.allclose(p.sum(1), 1.0)
    acc = (p.argmax(1) == y).mean()
    assert acc > 0.5


def test_lambdarank_gradients_match_oracle():
    X, rel, qb = datagen.rank_2m(400, 300)
    params = dict(objective="lambdarank", num_leaves=31, max_bin=63)
    sizes = np.diff(qb)
    ds = lgb.Dataset(X, label=rel, group=sizes, params=params)
    b = lgb.Booster(params, ds)
    rng = np.random.default_rng(13)
    s = np.round(rng.normal(size=len(rel)), 2)  # ties exercise the stable order
    g = np.zeros(len(rel), np.float32)
    h = np.zeros(len(rel), np.float32)
    lib = lgb._capi.lib()
    lgb._capi.check(lib.lgbmhip_gradients(b._ctx, lgb.ptr(s), lgb.ptr(g), lgb.ptr(h)), b._ctx)
    og, oh = O.lambdarank_gradients(s, rel, qb)
    # identical order of every f32 accumulation; only log2 of the normaliser may differ in its last ulp
    for a, c in ((g, og), (h, oh)):
        diff = a != c
        assert diff.sum() <= max(2, int(1e-4 * len(a)))
        if diff.any():
            assert np.all(np.abs(a[diff].view(np.int32) - c[diff].view(np.int32)) <= 1)


def test_lambdarank_boosting_improves_ndcg():
    X, rel, qb = datagen.rank_2m(300, 300)
    params = dict(objective="lambdarank", num_leaves=31, max_bin=63, min_data_in_leaf=20)
    ds = lgb.Dataset(X, label=rel, group=np.diff(qb), params=params)
    b = lgb.Booster(params, ds)
    ods = O.Dataset(X, rel, params)
    ods.set_query_boundaries(qb)
    ob = O.Booster(ods, params)
    n0 = O.ndcg(np.zeros(len(rel)), rel, qb, 10)
    done = _boost_both(b, ob, 5)
    n1 = O.ndcg(b.train_scores(), rel, qb, 10)
    n_cpu = O.ndcg(ob.scores(), rel, qb, 10)
    assert n1 > n0
    if done == 5:
        np.testing.assert_allclose(b.train_scores(), ob.scores(), rtol=1e-9, atol=1e-12)
        assert n1 == n_cpu
    else:
        assert abs(n1 - n_cpu) < 1e-3


def _categorical_data(n, seed):
    """binary_100k with column 1 replaced by a non-monotone category code of
    the informative column 24 (40 categories, 1% NaN); column 24 itself is
    shuffled, so its signal is reachable only through categorical splits."""
    X, y = datagen.binary_100k(n)
    rng = np.random.default_rng(seed)
    perm = rng.permutation(40)
    q = np.quantile(X[:, 24], np.linspace(0, 1, 41)[1:-1])
    X = X.copy()
    X[:, 1] = perm[np.digitize(X[:, 24], q)].astype(np.float64)
    X[rng.random(n) < 0.01, 1] = np.nan
    X[:, 24] = rng.permutation(X[:, 24])
    return X, y


def test_categorical_trees_and_prediction_match_oracle():
    X, y = _categorical_data(60000, 3)
    params = dict(P_BIN, num_leaves=63, max_bin=255, categorical_feature=[1], min_data_per_group=50,
                  cat_smooth=10.0)
    ds, b = _booster(X, y, params)
    ob = O.Booster(O.Dataset(X, y, params), params)
    for _ in range(4):
        b.update()
        ob.update()
    ncat = 0
    for t in range(4):
        gt, ot = b.tree(t), ob.tree(t)
        _compare_tree(gt, ot)
        ncat += int(np.sum(np.asarray(gt.decision_type) & 1))
    assert ncat > 0, "no categorical split was made"
    np.testing.assert_allclose(b.train_scores(), ob.scores(), rtol=1e-9, atol=1e-12)
    # raw-value prediction on training rows, unseen categories, negative
    # categories, NaN and zero (tree.rs:399-418 and predict.rs:95-109)
    Xt = np.concatenate([X[:5000], X[:200]])
    Xt[5000:5050, 1] = 57.0
    Xt[5050:5100, 1] = -3.0
    Xt[5100:5150, 1] = np.nan
    Xt[5150:5200, 0] = 0.0
    raw = b.predict(Xt, raw_score=True)
    np.testing.assert_allclose(raw, ob.predict_raw(Xt)[:, 0], rtol=1e-12, atol=1e-12)
    # binned prediction equals raw prediction on the training rows
    pb = b.predict_binned(ds.bins)[:, 0]
    r_all = b.predict(X, raw_score=True)
    assert np.all(np.abs(pb - r_all) <= 1e-9 * (1 + np.abs(r_all)))


def test_quantized_categorical_trees_and_prediction_match_oracle():
    X, y = _categorical_data(60000, 4)
    params = dict(P_BIN, num_leaves=63, max_bin=255, categorical_feature=[1], min_data_per_group=50,
                  cat_smooth=10.0, use_quantized_grad=True, num_grad_quant_bins=16, quant_train_renew_leaf=True)
    ds, b = _booster(X, y, params)
    ob = O.Booster(O.Dataset(X, y, params), params)
    for _ in range(3):
        b.update()
        ob.update()
    ncat = 0
    for t in range(3):
        gt = b.tree(t)
        _compare_tree(gt, ob.tree(t))
        ncat += int(np.sum(np.asarray(gt.decision_type) & 1))
    assert ncat > 0, "no categorical split was made"
    np.testing.assert_allclose(b.train_scores(), ob.scores(), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(b.predict(X[:3000], raw_score=True), ob.predict_raw(X[:3000])[:, 0],
                               rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("trunc", [30, 100])
def test_lambdarank_gradients_every_query_size_class(trunc):
    # queries of 1 to 2600 documents: sizes at and around powers of two, at
    # the kernel's shape boundaries (224 / 225 and 448 / 449: four, three and
    # two anchors per step in 896 staged terms; 512 / 513: arrays in LDS or in
    # the global scratch; 896 / 897: one or two partner chunks); ties in the
    # scores (two decimals), about 1% of scores -inf, and anchors past the
    # first 64 ranks (trunc 100)
    sizes = np.array([1, 2, 3, 63, 64, 65, 130, 224, 225, 255, 256, 257, 448, 449, 511, 512, 513, 700, 896, 897,
                      2047, 2048, 2049, 2600])
    qb = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    n = int(qb[-1])
    rng = np.random.default_rng(31 + trunc)
    X = rng.normal(size=(n, 6)).astype(np.float32)
    rel = rng.integers(0, 5, n).astype(np.float32)
    params = dict(objective="lambdarank", num_leaves=15, max_bin=63, lambdarank_truncation_level=trunc)
    ds = lgb.Dataset(X, label=rel, group=sizes, params=params)
    b = lgb.Booster(params, ds)
    s = np.round(rng.normal(size=n), 2)
    s[rng.random(n) < 0.01] = -np.inf
    g = np.zeros(n, np.float32)
    h = np.zeros(n, np.float32)
    lib = lgb._capi.lib()
    lgb._capi.check(lib.lgbmhip_gradients(b._ctx, lgb.ptr(s), lgb.ptr(g), lgb.ptr(h)), b._ctx)
    og, oh = O.lambdarank_gradients(s, rel, qb, trunc=trunc)
    for a, c in ((g, og), (h, oh)):
        diff = a != c
        assert diff.sum() <= max(2, int(1e-4 * len(a)))
        if diff.any():
            assert np.all(np.abs(a[diff].view(np.int32) - c[diff].view(np.int32)) <= 1)


@pytest.mark.parametrize("K", [1, 3])
def test_prediction_window_and_early_stop_match_oracle(K):
    rng = np.random.default_rng(21)
    n = 20000
    X = rng.normal(size=(n, 8)).astype(np.float32)
    X[:50, 2] = 1e-38          # below kZeroThreshold: predicted as 0
    if K == 1:
        y = (X[:, 0] + X[:, 1] > 0).astype(np.float32)
        params = dict(objective="binary", num_leaves=15, max_bin=63, learning_rate=0.3)
    else:
        y = np.argmax(X[:, :K] + 0.5 * rng.gumbel(size=(n, K)), 1).astype(np.float32)
        params = dict(objective="multiclass", num_class=K, num_leaves=15, max_bin=63, learning_rate=0.3)
    ds, b = _booster(X, y, params)
    ob = O.Booster(O.Dataset(X, y, params), params)
    for _ in range(12):
        b.update()
        ob.update()
    Xt = X[:4000]
    # trees equal the oracle's; leaf values within FP64 summation noise (1e-9)
    close = dict(rtol=1e-9, atol=1e-12)
    full = b.predict(Xt, raw_score=True).reshape(len(Xt), K)
    np.testing.assert_allclose(full, ob.predict_window(Xt), **close)
    win = b.predict(Xt, raw_score=True, start_iteration=2, num_iteration=5).reshape(len(Xt), K)
    np.testing.assert_allclose(win, ob.predict_window(Xt, 2, 5), **close)
    es = b.predict(Xt, raw_score=True, pred_early_stop=True, pred_early_stop_freq=3,
                   pred_early_stop_margin=1.0).reshape(len(Xt), K)
    ref = ob.predict_window(Xt, 0, -1, 3, 1.0)
    np.testing.assert_allclose(es, ref, **close)
    stopped = np.any(np.abs(es - full) > 1e-6, axis=1)
    assert 0 < stopped.sum() < len(Xt)   # some rows stopped early, not all


def test_leaf_prediction_and_rollback():
    X, y = datagen.binary_100k(30000)
    params = dict(P_BIN, num_leaves=31, max_bin=63)
    ds, b = _booster(X, y, params)
    dv = lgb.Dataset(X[:5000], label=y[:5000], reference=ds, params=params)
    b.add_valid(dv, "v")
    for _ in range(4):
        b.update()
    s3 = None
    # leaf indices: the training partition's leaf of each row for the last tree
    leaves = b.predict(X, pred_leaf=True)
    assert leaves.shape == (len(X), 4)
    np.testing.assert_array_equal(leaves[:, 3], b.leaf_of_row())
    np.testing.assert_array_equal(b.predict(X[:100], pred_leaf=True, start_iteration=1, num_iteration=2),
                                  leaves[:100, 1:3])
    # rollback: scores go back to those after 3 iterations (up to the f64 add/subtract rounding)
    b2 = lgb.Booster(params, ds)
    for _ in range(3):
        b2.update()
    s3 = b2.train_scores()
    b.rollback_one_iter()
    assert b.num_trees() == 3
    np.testing.assert_allclose(b.train_scores(), s3, rtol=0, atol=1e-12)
    np.testing.assert_allclose(b.valid_scores(0), b.predict_binned(dv.bins)[:, 0], rtol=0, atol=1e-12)
    # training continues from the rolled-back state
    b.update()
    assert b.num_trees() == 4
    imp = b.feature_importance()
    assert imp.sum() == sum(b.tree(t).num_leaves - 1 for t in range(4))


def test_lds_atomic_peak_is_a_rate():
    # the denominator of the histogram kernel's LDS-atomic fraction in bench
    # reports: lane atomics per second of conflict-free ds_add_u32
    X, y = small_binary(2000, 4)
    _, b = _booster(X, y, P_BIN)
    r = b.lds_atomic_peak()
    assert np.isfinite(r) and 1e11 < r < 1e15
