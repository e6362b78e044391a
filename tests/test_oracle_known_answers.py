"""Pins the CPU oracle against the reference's own known-answer tests and
against what the mathematics fixes (brute force, closed forms)."""
import math

import numpy as np
import pytest

from oracle import pyoracle as O


def test_mt19937_standard_draw():
    # mt19937.rs:95-103: the 10000th draw of mt19937(5489) is 4123659995
    assert O.lib().orc_mt19937_nth(5489, 10000) == 4123659995


def test_lcg_sample_properties():
    # random.rs:77-95
    out = np.zeros(10, np.int32)
    assert O.lib().orc_random_sample(1, 5, 5, np.zeros(5, np.int32)) == 5
    for n, k in [(1000, 10), (1000, 900), (50000, 20000)]:
        buf = np.zeros(k, np.int32)
        m = O.lib().orc_random_sample(7, n, k, buf)
        s = buf[:m]
        assert np.all(np.diff(s) > 0) and s.min() >= 0 and s.max() < n
        if k <= 10:
            assert m == k


def test_to_i8_casts_like_x86():
    # quant.rs:166-172
    assert O.lib().orc_to_i8(2.7) == 2
    assert O.lib().orc_to_i8(-2.7) == -2
    assert O.lib().orc_to_i8(200.0) == -56
    assert O.lib().orc_to_i8(float("nan")) == 0
    assert O.lib().orc_to_i8(1e300) == 0


def test_hist_fix_known_answers():
    # histogram.rs:91-98: bins 0,1,1,2 with gradients 1,2,3,4 (hessian 1
    # each) and most-frequent bin 1, which is not stored
    got = O.hist_fix(np.array([[1.0, 1.0], [0.0, 0.0], [4.0, 1.0]]), 1, 10.0, 4.0)
    assert got.tolist() == [1.0, 1.0, 5.0, 2.0, 4.0, 1.0]
    # histogram.rs:100-106: most-frequent bin 0 (offset 1) leaves the histogram alone
    got = O.hist_fix(np.array([[0.0, 0.0], [2.0, 1.0], [3.0, 1.0]]), 0, 7.0, 3.0)
    assert got.tolist() == [0.0, 0.0, 2.0, 1.0, 3.0, 1.0]


def _pk(g, h):
    return ((g << 32) + h + (1 << 63)) % (1 << 64) - (1 << 63)   # packed (g << 32) | u8(h) as int64


def test_hist_fix_int_on_the_same_rows():
    # the packed form of the rows above: bin 1 = total - bin 0 - bin 2
    got = O.hist_fix_int(np.array([_pk(1, 1), 0, _pk(4, 1)]), 1, _pk(10, 4))
    assert got.tolist() == [_pk(1, 1), _pk(5, 2), _pk(4, 1)]
    # gradients 1, -2, -3, 4: the total's gradient is 0, so bin 1 needs the
    # wrapping subtraction to come out as (-5 << 32) + 2
    got = O.hist_fix_int(np.array([_pk(1, 1), 0, _pk(4, 1)]), 1, _pk(0, 4))
    assert got.tolist() == [_pk(1, 1), _pk(-5, 2), _pk(4, 1)]
    got = O.hist_fix_int(np.array([0, _pk(2, 1), _pk(3, 1)]), 0, _pk(7, 3))
    assert got.tolist() == [0, _pk(2, 1), _pk(3, 1)]


def test_discretize_packing_known_answer():
    # quant.rs:175-184: bins 4, seed 0, no stochastic rounding
    gq, hq, gs, hs = O.discretize(np.array([1.0, -0.5], np.float32), np.array([0.5, 0.25], np.float32), 4, 0, False)
    assert list(hq) == [4, 2] and list(gq) == [2, -1]
    assert (gs, hs) == (0.5, 0.125)
    packed = [(int(g) << 32) | (int(h) & 0xFF) for g, h in zip(gq, hq)]
    assert packed[1] == (-1 << 32) | 2
    assert sum(packed) == (1 << 32) | 6
    sg = sum(float(g) * gs for g in gq)
    sh = sum(float(h) * hs for h in hq)
    assert (sg, sh) == (0.5, 0.75)


def test_binning_midpoints():
    # binning.rs:704-718
    vals = np.repeat(np.arange(1, 11, dtype=np.float64), 5)
    m = O.find_bin(vals, len(vals), max_bin=255, min_data_in_bin=3)
    kz = float(np.float32(1e-35))
    assert m["missing_type"] == 0
    assert m["upper"][0] == kz
    assert m["upper"][1] == np.nextafter(1.5, np.inf)
    assert m["upper"][-1] == np.inf
    vb = lambda v: O.lib().orc_value_to_bin_num(m["upper"], m["num_bin"], m["missing_type"], v)
    assert vb(1.0) == 1 and vb(10.0) == m["num_bin"] - 1 and vb(-5.0) == 0


def test_binning_nan_last_bin():
    # binning.rs:720-729
    vals = np.concatenate([np.arange(1, 51, dtype=np.float64), [np.nan, np.nan]])
    m = O.find_bin(vals, len(vals), max_bin=255, min_data_in_bin=3)
    assert m["missing_type"] == 2
    assert math.isnan(m["upper"][-1])
    assert O.lib().orc_value_to_bin_num(m["upper"], m["num_bin"], 2, float("nan")) == m["num_bin"] - 1


def test_binning_greedy_respects_max_bin():
    # binning.rs:731-739
    vals = np.sin(np.arange(10000) * 0.37) + 2.0
    m = O.find_bin(vals, len(vals), max_bin=16, min_data_in_bin=3)
    assert m["num_bin"] <= 16
    assert np.all(np.diff(m["upper"]) > 0)


def test_binning_categorical_order():
    # binning.rs:741-763: category 3 x30, 1 x20, 2 x10, -1 x5, zeros x15
    vals = np.concatenate([np.full(30, 3.0), np.full(20, 1.0), np.full(10, 2.7), np.full(5, -1.0)])
    m = O.find_bin(vals, len(vals) + 15, bin_type=1, max_bin=255, min_data_in_bin=3)
    assert list(m["bin_2_categorical"]) == [-1, 3, 1, 0, 2]
    assert m["missing_type"] == 2
    assert m["default_bin"] == 3


def test_binning_constant_is_trivial():
    m = O.find_bin(np.zeros(0), 100, max_bin=255, min_data_in_bin=3)
    assert m["is_trivial"]


def test_value_to_bin_matches_bruteforce():
    rng = np.random.default_rng(5)
    vals = rng.normal(size=3000)
    m = O.find_bin(vals, len(vals), max_bin=63, min_data_in_bin=3)
    up = m["upper"]
    for v in rng.normal(size=500):
        b = O.lib().orc_value_to_bin_num(up, m["num_bin"], m["missing_type"], float(v))
        # the first bin whose upper bound is >= v
        assert b == int(np.argmax(up >= v))


P1 = dict(lambda_l1=0.0, lambda_l2=0.0, max_delta_step=0.0, path_smooth=0.0, min_data_in_leaf=1,
          min_sum_hessian_in_leaf=1e-3, min_gain_to_split=0.0, max_cat_to_onehot=4, max_cat_threshold=32,
          cat_l2=10.0, cat_smooth=10.0, min_data_per_group=100)


def test_finds_obvious_split():
    # split.rs:1423-1443
    hist = np.array([[-5.0, 5.0], [-5.0, 5.0], [5.0, 5.0], [5.0, 5.0]])
    r = O.find_best_threshold(hist, 4, 0, 0, 0, 0, P1, 0.0, 20.0, 20)
    assert r["splittable"] and r["threshold"] == 1 and r["left_count"] == 10
    assert abs(r["left_output"] - 1.0) < 1e-12 and abs(r["right_output"] + 1.0) < 1e-12
    assert abs(r["gain"] - 20.0) < 1e-9


def test_categorical_one_hot_and_sorted():
    # split.rs:1445-1470
    hist = np.array([[0.0, 10.0], [-10.0, 10.0], [10.0, 10.0]])
    r = O.find_best_threshold(hist, 3, 2, 1, 1, 1, P1, 0.0, 30.0, 30)
    assert r["splittable"] and not r["default_left"] and r["left_count"] == 10
    assert r["cat_threshold"] == [1]
    # many-vs-many: 6 bins, offset 1 (bin 0 not stored); the two most negative ctr go left
    hist = np.array([[0.0, 0.0], [5.0, 100.0], [-50.0, 100.0], [2.0, 100.0], [-40.0, 100.0], [3.0, 100.0]])
    p = dict(P1, min_data_per_group=1, cat_smooth=1.0)
    r = O.find_best_threshold(hist, 6, 2, 0, 0, 1, p, -80.0, 600.0, 600)
    assert r["splittable"] and r["cat_threshold"] == [2, 4]


def _packed(g, h):
    return (np.asarray(g, np.int64) << 32) + np.asarray(h, np.int64)


def test_categorical_int_one_hot_and_sorted():
    # the same known answers on packed integer sums with unit scales
    # (split.rs:1194-1400 reduces to split.rs:446-670 there, kEpsilon aside)
    hist = _packed([0, -10, 10], [10, 10, 10])
    r = O.find_best_threshold_int(hist, 3, 2, 1, 1, P1, int(_packed(0, 30)), 1.0, 1.0, 30, bin_type=1)
    assert r["splittable"] and not r["default_left"] and r["left_count"] == 10
    assert r["cat_threshold"] == [1]
    assert r["left_sum_gh"] == int(_packed(-10, 10)) and r["right_sum_gh"] == int(_packed(10, 20))
    hist = _packed([0, 5, -50, 2, -40, 3], [0, 100, 100, 100, 100, 100])
    p = dict(P1, min_data_per_group=1, cat_smooth=1.0)
    r = O.find_best_threshold_int(hist, 6, 2, 0, 0, p, int(_packed(-80, 600)), 1.0, 1.0, 600, bin_type=1)
    assert r["splittable"] and r["cat_threshold"] == [2, 4]
    assert r["left_sum_gh"] == int(_packed(-90, 200)) and r["left_count"] == 200


def test_categorical_int_agrees_with_float_search():
    """Integer-valued histograms with unit scales and count = hessian: the
    packed-sum search and the FP64 search must choose the same category set,
    counts and sums, and gains/outputs equal up to the FP64 form's kEpsilon."""
    rng = np.random.default_rng(21)
    agree = 0
    for trial in range(300):
        nb = int(rng.integers(2, 48))
        onehot = nb <= 4
        cnt = rng.integers(0, 80, nb)
        g = (rng.normal(size=nb) * cnt * rng.uniform(0.2, 2.0, nb)).astype(np.int64)
        mfb = int(rng.integers(0, nb))
        tot_g, tot_h, nd = int(g.sum()), int(cnt.sum()), int(cnt.sum())
        if nd == 0:
            continue
        p = dict(P1, min_data_in_leaf=int(rng.integers(1, 20)), min_data_per_group=int(rng.integers(1, 60)),
                 cat_smooth=float(rng.choice([1.0, 5.0, 10.0])), max_cat_threshold=int(rng.integers(1, 32)))
        hf = np.stack([g, cnt], 1).astype(np.float64)
        hf[mfb] = 0.0
        hi = _packed(g, cnt)
        hi[mfb] = 0
        rf = O.find_best_threshold(hf, nb, 2, mfb, mfb, 1, p, float(tot_g), float(tot_h), nd)
        ri = O.find_best_threshold_int(hi, nb, 2, mfb, mfb, p, int(_packed(tot_g, tot_h)), 1.0, 1.0, nd, bin_type=1)
        assert rf["splittable"] == ri["splittable"], trial
        if not ri["splittable"]:
            continue
        assert sorted(rf["cat_threshold"]) == sorted(ri["cat_threshold"]), (trial, onehot)
        assert (rf["left_count"], rf["right_count"]) == (ri["left_count"], ri["right_count"])
        # the chosen set's packed sum is the wrapping sum of its bins (mfb restored from the total)
        full = hi.copy()
        full[mfb] = int(_packed(tot_g, tot_h)) - int(np.delete(hi, mfb).sum())
        assert ri["left_sum_gh"] == int(full[ri["cat_threshold"]].sum())
        assert ri["left_sum_gh"] + ri["right_sum_gh"] == int(_packed(tot_g, tot_h))
        assert abs(rf["gain"] - ri["gain"]) <= 1e-9 * max(1.0, abs(ri["gain"]))
        assert abs(rf["left_output"] - ri["left_output"]) <= 1e-9 * max(1.0, abs(ri["left_output"]))
        agree += 1
    assert agree >= 100


def _bruteforce_best(hist, sg, sh, nd, p):
    """All thresholds of a numerical feature without missing values (mfb = 0
    so offset = 1 and bin 0 is the remainder), scored by G^2/H."""
    eps = float(np.float32(1e-15))
    shp = sh + 2 * eps
    cf = nd / shp
    best = (-np.inf, None)
    B = len(hist)
    for thr in range(0, B - 1):
        rg = hist[thr + 1:, 0].sum()
        rh = eps + hist[thr + 1:, 1].sum()
        rc = sum(int(h * cf + 0.5) for h in hist[thr + 1:, 1])
        lc = nd - rc
        lh = shp - rh
        lg = sg - rg
        if rc < p["min_data_in_leaf"] or lc < p["min_data_in_leaf"]:
            continue
        gain = lg * lg / lh + rg * rg / rh - sg * sg / shp
        if gain > best[0] + 1e-9:
            best = (gain, thr)
    return best


def test_threshold_search_matches_bruteforce():
    rng = np.random.default_rng(11)
    p = dict(P1, min_data_in_leaf=3)
    for _ in range(30):
        B = int(rng.integers(3, 20))
        cnt = rng.integers(1, 30, size=B).astype(np.float64)
        g = rng.normal(size=B) * cnt
        hist = np.stack([g, cnt], axis=1)
        hist[0] = 0.0  # bin 0 is the most frequent bin (not stored)
        tot_g = g.sum()
        tot_h = cnt.sum()
        nd = int(cnt.sum())
        gain, thr = _bruteforce_best(np.stack([np.r_[tot_g - g[1:].sum(), g[1:]], cnt], 1), tot_g, tot_h, nd, p)
        r = O.find_best_threshold(hist, B, 0, 0, 0, 0, p, tot_g, tot_h, nd)
        if thr is None:
            assert not r["splittable"] or r["gain"] <= 1e-9
        else:
            assert r["threshold"] == thr
            assert abs(r["gain"] - gain) <= 1e-6 * max(1.0, abs(gain))


def test_tie_break_prefers_smaller_feature_in_training():
    # two identical columns: the split must use the smaller column index
    rng = np.random.default_rng(1)
    n = 500
    x = rng.normal(size=n).astype(np.float32)
    X = np.stack([x, x], axis=1)
    y = (x > 0.1).astype(np.float32)
    ds = O.Dataset(X, y, dict(max_bin=63, min_data_in_leaf=5))
    b = O.Booster(ds, dict(objective="binary", num_leaves=4, min_data_in_leaf=5, max_bin=63))
    b.update()
    t = b.tree(0)
    assert set(t.split_feature.tolist()) == {0}


def _step_data(n):
    a = np.arange(n) / n
    X = np.stack([a, ((np.arange(n) * 37) % 11).astype(np.float64)], axis=1).astype(np.float32)
    y = np.where(a < 0.5, -1.0, 1.0).astype(np.float32)
    return X, y


def test_learner_first_split_separates_step():
    # learner/tests.rs:24-43
    n = 200
    X, y = _step_data(n)
    ds = O.Dataset(X, y, dict(min_data_in_leaf=5))
    b = O.Booster(ds, dict(objective="binary", num_leaves=2, min_data_in_leaf=5))
    b.train_tree(-y, np.ones(n, np.float32))
    t = b.tree(0)
    assert t.num_leaves == 2 and t.split_feature[0] == 0
    assert list(t.leaf_count) == [100, 100]
    assert abs(t.leaf_value[0] + 1.0) < 1e-9 and abs(t.leaf_value[1] - 1.0) < 1e-9


def test_learner_partition_stable_and_max_depth():
    # learner/tests.rs:45-78
    n = 300
    X, y = _step_data(n)
    ds = O.Dataset(X, y, dict(min_data_in_leaf=40))
    b = O.Booster(ds, dict(objective="binary", num_leaves=31, max_depth=2, min_data_in_leaf=40))
    b.train_tree(np.sin(np.arange(n, dtype=np.float32) * 0.37).astype(np.float32), np.ones(n, np.float32))
    t = b.tree(0)
    assert t.num_leaves <= 4
    assert np.all(t.leaf_count >= 40)


def test_learner_constant_gradient_single_leaf():
    # learner/tests.rs:80-89
    n = 100
    X, y = _step_data(n)
    ds = O.Dataset(X, y, {})
    b = O.Booster(ds, dict(objective="binary"))
    assert b.train_tree(np.full(n, 0.5, np.float32), np.ones(n, np.float32)) == 1


def test_binary_gradients_closed_form():
    rng = np.random.default_rng(2)
    s = rng.normal(size=1000)
    y = (rng.random(1000) > 0.5).astype(np.float32)
    g, h = O.binary_gradients(s, y)
    lab = np.where(y > 0, 1.0, -1.0)
    resp = -lab / (1.0 + np.exp(lab * s))
    np.testing.assert_array_equal(g, resp.astype(np.float32))
    np.testing.assert_array_equal(h, (np.abs(resp) * (1.0 - np.abs(resp))).astype(np.float32))


def test_multiclass_gradients_closed_form():
    rng = np.random.default_rng(3)
    K, n = 4, 500
    s = rng.normal(size=K * n)
    y = rng.integers(0, K, size=n).astype(np.float32)
    g, h = O.multiclass_gradients(s, y, K)
    S = s.reshape(K, n).T
    P = np.exp(S - S.max(1, keepdims=True))
    P = P / P.sum(1, keepdims=True)
    onehot = np.eye(K)[y.astype(int)]
    np.testing.assert_allclose(g.reshape(K, n).T, P - onehot, atol=1e-6)
    np.testing.assert_allclose(h.reshape(K, n).T, K / (K - 1.0) * P * (1 - P), atol=1e-6)


def test_gradients_are_finite_differences_of_the_losses():
    # objective/tests.rs:18-56: g = dL/ds, h = d2L/ds2 by central differences
    # of the loss itself (binary logloss; softmax cross-entropy, whose
    # hessian the reference scales by K / (K - 1), multiclass.rs:72)
    rng = np.random.default_rng(5)
    s = rng.normal(size=200) * 2
    y = (rng.random(200) > 0.5).astype(np.float32)
    g, h = O.binary_gradients(s, y)
    lab = np.where(y > 0, 1.0, -1.0)
    loss = lambda t: np.log1p(np.exp(-lab * t))
    e = 1e-4
    np.testing.assert_allclose(g, (loss(s + e) - loss(s - e)) / (2 * e), rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(h, (loss(s + e) - 2 * loss(s) + loss(s - e)) / e ** 2, rtol=1e-4, atol=1e-6)
    K, n = 3, 100
    s = rng.normal(size=K * n)
    y = rng.integers(0, K, size=n)
    g, h = O.multiclass_gradients(s, y.astype(np.float32), K)
    S = s.reshape(K, n).T

    def ce(S):
        m = S.max(1, keepdims=True)
        return -(S[np.arange(n), y] - m[:, 0] - np.log(np.exp(S - m).sum(1)))
    for c in range(K):
        d = np.zeros_like(S)
        d[:, c] = e
        fd1 = (ce(S + d) - ce(S - d)) / (2 * e)
        fd2 = (ce(S + d) - 2 * ce(S) + ce(S - d)) / e ** 2
        np.testing.assert_allclose(g.reshape(K, n)[c], fd1, rtol=1e-5, atol=1e-7)
        np.testing.assert_allclose(h.reshape(K, n)[c] * (K - 1.0) / K, fd2, rtol=1e-4, atol=1e-6)


def test_auc_matches_sklearn():
    from sklearn.metrics import roc_auc_score
    rng = np.random.default_rng(4)
    y = (rng.random(2000) > 0.6).astype(np.float32)
    s = rng.normal(size=2000) + y
    assert abs(O.auc(s, y) - roc_auc_score(y, s)) < 1e-12
    s2 = np.round(s, 1)  # ties
    assert abs(O.auc(s2, y) - roc_auc_score(y, s2)) < 1e-12


def test_training_scores_equal_raw_prediction():
    # end_to_end.rs:53-57
    rng = np.random.default_rng(9)
    n, p = 2000, 5
    X = rng.uniform(-2, 2, size=(n, p)).astype(np.float32)
    X[::7, 2] = np.nan
    X[::3, 3] = 0.0
    y = (X[:, 0] - 0.5 * X[:, 1] ** 2 + np.nan_to_num(X[:, 2], nan=1.0) > 0).astype(np.float32)
    ds = O.Dataset(X, y, dict(max_bin=255))
    b = O.Booster(ds, dict(objective="binary", num_leaves=15, learning_rate=0.1))
    for _ in range(20):
        b.update()
    raw = b.predict_raw(X)[:, 0]
    s = b.scores()
    assert np.all(np.abs(raw - s) <= 1e-9 * (1 + np.abs(s)))
    assert O.binary_logloss(s, y) < 0.4


def test_dcg_known_answers():
    # dcg.rs:195-209
    label = np.array([0.0, 2.0, 1.0], np.float32)
    ideal = 3.0 + 1.0 / np.log2(3.0)
    assert O.max_dcg_at_k(label, 3) == ideal
    assert O.max_dcg_at_k(label, 1) == 3.0


def test_lambdarank_pair_known_answer():
    # objective/rank.rs:491-509: one query, two documents, no normalisation
    g, h = O.lambdarank_gradients(np.zeros(2), np.array([1.0, 0.0], np.float32), np.array([0, 2]), norm=False)
    dndcg = 1.0 * (1.0 - 1.0 / np.log2(3.0))
    # the table value at delta 0 is sigmoid of the bin's score, within 1e-4 of 0.5
    p = 0.5
    assert abs(float(-g[0]) / dndcg - p) < 1e-4
    assert g[0] == -g[1] and g[0] < 0
    assert h[0] == h[1] and abs(float(h[0]) / dndcg - 0.25) < 1e-4


def test_lambdarank_bruteforce_small_queries():
    """Pairwise LambdaRank restated in numpy (with the reference's sigmoid table
    and f32 accumulation order) on tiny queries."""
    rng = np.random.default_rng(21)
    sizes = [1, 2, 5, 9, 40]
    qb = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    n = qb[-1]
    lab = rng.integers(0, 5, n).astype(np.float32)
    s = np.round(rng.normal(size=n), 1)  # ties exercise the stable order
    g, h = O.lambdarank_gradients(s, lab, qb, norm=True, trunc=30)
    bins = 1024 * 1024
    mn = -50.0 / 2.0
    fac = bins / (2 * -mn)
    gain = np.array([0.0] + [(1 << i) - 1 for i in range(1, 31)], np.float64)
    disc = 1.0 / np.log2(2.0 + np.arange(10000))

    def sig(x):
        if x <= mn:
            i = 0
        elif x >= -mn:
            i = bins - 1
        else:
            i = int((x - mn) * fac)
        return 1.0 / (1.0 + np.exp((i / fac + mn)))

    for q in range(len(sizes)):
        a, b = qb[q], qb[q + 1]
        cnt = b - a
        if cnt == 0:
            continue
        ls, ll = s[a:b], lab[a:b]
        order = sorted(range(cnt), key=lambda i: -ls[i])  # Python's sort is stable
        srt = sorted(ll, reverse=True)
        mx = sum(disc[j] * gain[int(srt[j])] for j in range(min(30, cnt)))
        imd = 1.0 / mx if mx > 0 else mx
        lam = np.zeros(cnt, np.float32)
        hes = np.zeros(cnt, np.float32)
        best, worst = ls[order[0]], ls[order[-1]]
        S = 0.0
        for i in range(min(cnt - 1, 30)):
            for j in range(i + 1, cnt):
                if ll[order[i]] == ll[order[j]]:
                    continue
                hr, lr = (i, j) if ll[order[i]] > ll[order[j]] else (j, i)
                hi, lo = order[hr], order[lr]
                ds = ls[hi] - ls[lo]
                dp = (gain[int(ll[hi])] - gain[int(ll[lo])]) * abs(disc[hr] - disc[lr]) * imd
                if best != worst:
                    dp /= float(np.float32(0.01)) + abs(ds)
                pl = sig(ds)
                ph = pl * (1.0 - pl)
                pl *= -1.0 * dp
                ph *= 1.0 * 1.0 * dp
                lam[lo] = np.float32(lam[lo] - np.float32(pl))
                hes[lo] = np.float32(hes[lo] + np.float32(ph))
                lam[hi] = np.float32(lam[hi] + np.float32(pl))
                hes[hi] = np.float32(hes[hi] + np.float32(ph))
                S -= 2.0 * pl
        if S > 0:
            f = np.log2(1.0 + S) / S
            lam = (lam.astype(np.float64) * f).astype(np.float32)
            hes = (hes.astype(np.float64) * f).astype(np.float32)
        np.testing.assert_array_equal(g[a:b], lam)
        np.testing.assert_array_equal(h[a:b], hes)


def test_ndcg_matches_sklearn_on_tie_free_queries():
    from sklearn.metrics import ndcg_score
    rng = np.random.default_rng(22)
    sizes = [5, 12, 30, 7]
    qb = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    lab = rng.integers(0, 5, qb[-1]).astype(np.float32)
    s = rng.normal(size=qb[-1])
    ours = O.ndcg(s, lab, qb, 10)
    ref = []
    for q in range(len(sizes)):
        a, b = qb[q], qb[q + 1]
        if lab[a:b].max() == 0:
            ref.append(1.0)
            continue
        ref.append(ndcg_score([2.0 ** lab[a:b] - 1], [s[a:b]], k=10))
    assert abs(ours - np.mean(ref)) < 1e-12


# ---------------------------------------------------------------- metrics (metric.rs:551-729)
def test_metric_auc_matches_sklearn_and_pairs():
    from sklearn.metrics import roc_auc_score
    rng = np.random.default_rng(30)
    for n, ties in ((500, False), (3000, True)):
        s = rng.normal(size=n)
        if ties:
            s = np.round(s, 1)
        y = (rng.random(n) < 0.35).astype(np.float32)
        w = (rng.random(n) * 3).astype(np.float32)
        assert O.metric_auc(s, y) == pytest.approx(roc_auc_score(y, s), rel=1e-13)
        assert O.metric_auc(s, y, w) == pytest.approx(roc_auc_score(y, s, sample_weight=w), rel=1e-12)
    # brute force over (positive, negative) pairs, ties count one half
    s = np.round(rng.normal(size=60), 1)
    y = (rng.random(60) < 0.5).astype(np.float32)
    w = rng.random(60).astype(np.float32)
    pos, neg = np.where(y > 0)[0], np.where(y <= 0)[0]
    num = sum(float(w[p]) * float(w[q]) * (1.0 if s[p] > s[q] else 0.5 if s[p] == s[q] else 0.0)
              for p in pos for q in neg)
    den = float(w[pos].astype(np.float64).sum()) * float(w[neg].astype(np.float64).sum())
    assert O.metric_auc(s, y, w) == pytest.approx(num / den, rel=1e-12)
    # one class only: 1 (metric.rs:724-728)
    assert O.metric_auc(s, np.ones(60, np.float32)) == 1.0
    # -0.0 and +0.0 are one score
    assert O.metric_auc(np.array([0.0, -0.0]), np.array([1, 0], np.float32)) == 0.5


def test_metric_point_matches_sklearn():
    from sklearn.metrics import log_loss
    rng = np.random.default_rng(31)
    n = 4000
    s = rng.normal(size=n) * 2
    y = (rng.random(n) < 0.5).astype(np.float32)
    w = (rng.random(n) + 0.2).astype(np.float32)
    p = 1.0 / (1.0 + np.exp(-s))
    assert O.metric_point("binary_logloss", s, y, w) == pytest.approx(log_loss(y, p, sample_weight=w), rel=1e-12)
    err = np.where(p <= 0.5, y > 0, y <= 0).astype(np.float64)
    assert O.metric_point("binary_error", s, y) == err.mean()
    # the clamp: -log(float eps) when the probability underflows
    big = O.metric_point("binary_logloss", np.array([800.0]), np.array([0.0], np.float32))
    assert big == -float(np.log(np.float32(1e-15)))
    K = 4
    S = rng.normal(size=(K, n))
    yk = rng.integers(0, K, n).astype(np.float32)
    P = np.exp(S - S.max(0)) / np.exp(S - S.max(0)).sum(0)
    assert O.metric_point("multi_logloss", S, yk, K=K) == pytest.approx(
        log_loss(yk, P.T, labels=list(range(K))), rel=1e-12)
    rank = (P >= P[yk.astype(int), np.arange(n)]).sum(0)
    for top_k in (1, 2, 3):
        assert O.metric_point("multi_error", S, yk, K=K, top_k=top_k) == (rank > top_k).mean()


def test_metric_ndcg_known_answers_and_sklearn():
    from sklearn.metrics import ndcg_score
    # dcg.rs:195-209: labels [0, 2, 1], scores [1, 0, 0] (the tie keeps index order)
    lab = np.array([0, 2, 1], np.float32)
    out = O.metric_ndcg(np.array([1.0, 0.0, 0.0]), lab, [0, 3], (1, 3))
    ideal = 3.0 + 1.0 / np.log2(3.0)
    assert out[0] == 0.0
    assert out[1] == pytest.approx((3.0 / np.log2(3.0) + 1.0 / np.log2(4.0)) / ideal, rel=1e-15)
    # a query without relevant documents counts 1
    assert O.metric_ndcg(np.zeros(4), np.zeros(4, np.float32), [0, 4], (1, 2)).tolist() == [1.0, 1.0]
    rng = np.random.default_rng(32)
    sizes = [5, 12, 30, 7, 50]
    qb = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    lab = rng.integers(0, 5, qb[-1]).astype(np.float32)
    lab[qb[1]] = 4
    s = rng.normal(size=qb[-1])
    at = (1, 3, 10)
    ours = O.metric_ndcg(s, lab, qb, at)
    for m, k in enumerate(at):
        ref = [ndcg_score([2.0 ** lab[a:b] - 1], [s[a:b]], k=k) if lab[a:b].max() > 0 else 1.0
               for a, b in zip(qb[:-1], qb[1:])]
        assert ours[m] == pytest.approx(np.mean(ref), rel=1e-12)
    # per-thread partial sums change only the rounding
    assert O.metric_ndcg(s, lab, qb, at, threads=3) == pytest.approx(ours, rel=1e-14)


# ---------------------------------------------------------------- bagging (sample_strategy.rs:186-318)
def _oracle_bag(n, seed, fraction):
    X = np.random.default_rng(0).normal(size=(n, 3)).astype(np.float32)
    y = (X[:, 0] > 0).astype(np.float32)
    p = dict(objective="binary", num_leaves=4, bagging_fraction=fraction, bagging_freq=1, bagging_seed=seed)
    ob = O.Booster(O.Dataset(X, y, p), p)
    ob.update()
    return ob.bag(), ob


def test_bagging_draw_is_seeded_per_1024_row_block():
    n = 5 * 1024
    a, _ = _oracle_bag(n, 7, 0.37)
    b, _ = _oracle_bag(n, 8, 0.37)
    # block j's generator is Random(seed + j): seed 8's block j draws what seed 7's block j + 1 draws
    for j in range(4):
        aj = a[(a >= (j + 1) * 1024) & (a < (j + 2) * 1024)] - 1024
        bj = b[(b >= j * 1024) & (b < (j + 1) * 1024)]
        np.testing.assert_array_equal(aj, bj)
    assert np.all(np.diff(a) > 0)
    assert abs(len(a) / n - 0.37) < 0.03
    # the first draw of a block is Random(seed + j)'s first float (random.rs:19-33)
    for j in range(5):
        x = ((7 + j) * 214013 + 2531011) & 0xFFFFFFFF
        first = np.float32(((x >> 16) & 0x7FFF) / np.float32(32768.0))
        assert (j * 1024 in set(a.tolist())) == (float(first) < 0.37)


def test_bagging_generators_run_on_between_draws():
    n = 3000
    X = np.random.default_rng(1).normal(size=(n, 3)).astype(np.float32)
    y = (X[:, 0] > 0).astype(np.float32)
    p = dict(objective="binary", num_leaves=4, bagging_fraction=0.5, bagging_freq=1, bagging_seed=3)
    ob = O.Booster(O.Dataset(X, y, p), p)
    ob.update()
    first = ob.bag().copy()
    ob.update()
    second = ob.bag()
    assert not np.array_equal(first, second)


def _lcg_streams(seeds, count):
    """count draws of Random(seed) for each seed (random.rs:19-33, 97-105):
    x' = 214013 x + 2531011 (mod 2^32), next_float = ((x' >> 16) & 0x7FFF) / 32768 in f32."""
    x = np.array(seeds, np.uint64) & np.uint64(0xFFFFFFFF)
    out = np.empty((len(seeds), count), np.float32)
    for i in range(count):
        x = (x * np.uint64(214013) + np.uint64(2531011)) & np.uint64(0xFFFFFFFF)
        out[:, i] = ((x >> np.uint64(16)) & np.uint64(0x7FFF)).astype(np.float32) / np.float32(32768.0)
    return out


def _expected_bags(n, seed, fracs_of_row, bags):
    """The bags of `bags` consecutive draws: block j's generator is
    Random(seed + j) and runs on from one bag to the next; row i is in the bag
    when its draw (as f64) is below its fraction (sample_strategy.rs:285-315)."""
    nb = (n + 1023) // 1024
    streams = _lcg_streams([seed + j for j in range(nb)], 1024 * bags)
    out = []
    for k in range(bags):
        draws = np.empty(n, np.float32)
        for j in range(nb):
            lo, hi = j * 1024, min(n, (j + 1) * 1024)
            draws[lo:hi] = streams[j, k * (hi - lo):(k + 1) * (hi - lo)]
        out.append(np.flatnonzero(draws.astype(np.float64) < fracs_of_row))
    return out


@pytest.mark.parametrize("balanced", [False, True])
def test_bag_draws_follow_the_lcg_streams(balanced):
    n = 20 * 1024 + 300   # a ragged last block
    rng = np.random.default_rng(21)
    X = rng.normal(size=(n, 3)).astype(np.float32)
    y = (rng.random(n) < 0.35).astype(np.float32)
    p = dict(objective="binary", num_leaves=4, bagging_freq=1, bagging_seed=5)
    if balanced:
        p.update(pos_bagging_fraction=0.3, neg_bagging_fraction=0.6)
        fr = np.where(y > 0, 0.3, 0.6)
    else:
        p.update(bagging_fraction=0.45)
        fr = np.full(n, 0.45)
    ob = O.Booster(O.Dataset(X, y, p), p)
    want = _expected_bags(n, 5, fr, 3)
    for k in range(3):
        ob.update()
        np.testing.assert_array_equal(ob.bag(), want[k])


def test_goss_block_info_known_answers():
    # sample_strategy.rs:457-463 (block_info_force_size)
    assert O.block_info_force_size(1, 5000, 1024) == (1, 5000)
    assert O.block_info_force_size(4, 5000, 1024) == (4, 2048)
    assert O.block_info_force_size(8, 3000, 1024) == (3, 1024)
    assert O.block_info_force_size(4, 500, 1024) == (1, 500)


def _check_goss_block_properties(v, g0, h0, g1, h1, bag, n, K, threads, top_rate, other_rate):
    """Per sampling block of block_info_force_size(threads, n, 1024): every row
    whose |g h| reaches the block's top_k-th largest value is kept; at most
    other_k other rows are sampled, exactly other_k when no value ties the
    threshold (selection sampling with probability rest_need / rest_all keeps
    exactly that many); sampled rows' g and h are multiplied by
    f32(cnt - top_k) / f32(other_k), every other row's are untouched."""
    in_bag = np.zeros(n, bool)
    in_bag[bag] = True
    assert np.all(np.diff(bag.astype(np.int64)) > 0)
    nblock, block = O.block_info_force_size(threads, n, 1024)
    scaled = np.zeros(n, bool)
    exact_blocks = 0
    for b in range(nblock):
        lo = b * block
        if lo >= n:
            break
        hi = min(n, lo + block)
        cnt = hi - lo
        top_k = max(int(cnt * top_rate), 1)
        other_k = int(cnt * other_rate)
        vb = v[lo:hi]
        thr = np.sort(vb)[::-1][top_k - 1]
        top = vb >= thr
        assert in_bag[lo:hi][top].all()
        small_in = in_bag[lo:hi] & ~top
        if top.sum() == top_k:
            assert small_in.sum() == other_k
            exact_blocks += 1
        else:
            assert small_in.sum() <= other_k
        scaled[lo:hi] = small_in
        mult = np.float32(cnt - top_k) / np.float32(other_k)
        for k in range(K):
            np.testing.assert_array_equal(g1[k, lo:hi][small_in], g0[k, lo:hi][small_in] * mult)
            np.testing.assert_array_equal(h1[k, lo:hi][small_in], h0[k, lo:hi][small_in] * mult)
    for k in range(K):
        np.testing.assert_array_equal(g1[k][~scaled], g0[k][~scaled])
        np.testing.assert_array_equal(h1[k][~scaled], h0[k][~scaled])
    return exact_blocks


def _abs_gh(g, h, K):
    v = np.zeros(g.shape[1], np.float32)
    for k in range(K):   # f32 sum over the classes in order (sample_strategy.rs:404-411)
        v = v + np.abs(g[k] * h[k])
    return v


@pytest.mark.parametrize("K,threads", [(1, 3), (1, 1), (3, 2), (2, 8)])
def test_goss_sample_properties_on_continuous_gradients(K, threads):
    n = 9 * 1024 + 77
    rng = np.random.default_rng(30 + K + threads)
    g0 = rng.normal(size=(K, n)).astype(np.float32)
    h0 = (0.05 + rng.random(size=(K, n))).astype(np.float32)
    bag, g1, h1 = O.goss(g0, h0, K, threads, 0.2, 0.15, 7)
    g1, h1 = g1.reshape(K, n), h1.reshape(K, n)
    exact = _check_goss_block_properties(_abs_gh(g0, h0, K), g0, h0, g1, h1, bag, n, K, threads, 0.2, 0.15)
    nblock, block = O.block_info_force_size(threads, n, 1024)
    assert exact == sum(1 for b in range(nblock) if b * block < n)   # no ties: exactly other_k in every block


@pytest.mark.parametrize("K,threads", [(1, 3), (3, 2)])
def test_goss_sample_properties_in_boosting(K, threads):
    # boosting scores take few distinct values, so |g h| ties at the threshold
    n = 6 * 1024 + 100
    rng = np.random.default_rng(22 + K)
    X = rng.normal(size=(n, 4)).astype(np.float32)
    if K == 1:
        y = (X[:, 0] + 0.5 * rng.normal(size=n) > 0).astype(np.float32)
        p = dict(objective="binary")
    else:
        y = np.argmax(X[:, :K] + rng.gumbel(size=(n, K)), 1).astype(np.float32)
        p = dict(objective="multiclass", num_class=K)
    p.update(num_leaves=7, learning_rate=0.5, data_sample_strategy="goss", top_rate=0.2, other_rate=0.15,
             num_threads=threads, bagging_seed=3)
    ob = O.Booster(O.Dataset(X, y, p), p)
    for _ in range(2):   # GOSS starts at iteration 1 / learning_rate = 2
        ob.update()
    scores = ob.scores()
    ob.update()
    g1, h1 = ob.gradients()
    g0, h0 = O.binary_gradients(scores, y) if K == 1 else O.multiclass_gradients(scores, y, K)
    g0, h0, g1, h1 = (a.reshape(K, n) for a in (g0, h0, g1, h1))
    _check_goss_block_properties(_abs_gh(g0, h0, K), g0, h0, g1, h1, ob.bag(), n, K, threads, 0.2, 0.15)


@pytest.mark.parametrize("form", ["plain", "quant", "multiclass_cat"])
def test_oracle_threads_do_not_change_trees_or_scores(form):
    """The threaded CPU engine (histograms and searches by feature, stable
    partition by row chunks, per-row and per-leaf loops) gives the
    single-thread trees and scores bit for bit."""
    rng = np.random.default_rng(40)
    n = 150000
    X = rng.normal(size=(n, 10)).astype(np.float32)
    p = dict(objective="binary", num_leaves=63, max_bin=255, min_data_in_leaf=20, learning_rate=0.1)
    if form == "quant":
        p.update(use_quantized_grad=True, num_grad_quant_bins=16, quant_train_renew_leaf=True)
    if form == "multiclass_cat":
        X[:, 3] = rng.integers(0, 30, n)
        y = np.argmax(X[:, :3] + 0.1 * (X[:, 3:4] % 7) + rng.gumbel(size=(n, 3)), 1).astype(np.float32)
        p.update(objective="multiclass", num_class=3, categorical_feature=[3])
    else:
        y = (X[:, 0] + 0.5 * X[:, 1] * X[:, 2] + rng.normal(size=n) > 0).astype(np.float32)
    runs = []
    try:
        for threads in (1, 4):
            O.set_num_threads(threads)
            ob = O.Booster(O.Dataset(X, y, p), p)
            for _ in range(3):
                ob.update()
            runs.append((ob.scores(), [ob.tree(t) for t in range(ob.num_trees())]))
    finally:
        O.set_num_threads(1)
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    for a, b in zip(runs[0][1], runs[1][1]):
        for f in ("split_feature", "threshold_in_bin", "left_child", "right_child", "leaf_value", "leaf_count"):
            np.testing.assert_array_equal(getattr(a, f), getattr(b, f))


def test_bagging_by_query_draws_whole_queries_from_the_lcg_streams():
    """sample_strategy.rs:239-276: query q is drawn by the generator of its
    1024-query block (Random(seed + q // 1024), running on from bag to bag)
    and the bag is the rows of the drawn queries, in order."""
    rng = np.random.default_rng(23)
    sizes = rng.integers(1, 12, 2600)           # 3 query blocks, the last one ragged
    qb = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    n = int(qb[-1])
    X = rng.normal(size=(n, 3)).astype(np.float32)
    y = rng.integers(0, 4, n).astype(np.float32)
    p = dict(objective="lambdarank", num_leaves=4, bagging_fraction=0.4, bagging_freq=1, bagging_seed=9,
             bagging_by_query=True)
    ods = O.Dataset(X, y, p)
    ods.set_query_boundaries(qb)
    ob = O.Booster(ods, p)
    nq = len(sizes)
    nqb = (nq + 1023) // 1024
    streams = _lcg_streams([9 + j for j in range(nqb)], 1024 * 3)
    for k in range(3):
        ob.update()
        drawn = np.empty(nq, np.float32)
        for j in range(nqb):
            lo, hi = j * 1024, min(nq, (j + 1) * 1024)
            drawn[lo:hi] = streams[j, k * (hi - lo):(k + 1) * (hi - lo)]
        sampled = np.flatnonzero(drawn.astype(np.float64) < 0.4)
        want = np.concatenate([np.arange(qb[q], qb[q + 1]) for q in sampled])
        np.testing.assert_array_equal(ob.bag(), want)


@pytest.mark.parametrize("fraction", [0.3, 0.8])
def test_quantized_bag_root_sums(fraction):
    """learner/mod.rs:611-628 with a bag: in subset mode (average bag rate
    <= 0.5) the quantizer runs over the bag and the root sums are the bag
    rows' dequantized values and packed sums; with a full-data bag the
    quantizer runs over every row and the sums are quant.rs:144-158's
    bag_root_sums: f32 scales, and the packed sum of positions 0 .. m - 1."""
    rng = np.random.default_rng(24)
    n = 30000
    X = rng.normal(size=(n, 5)).astype(np.float32)
    y = (X[:, 0] + rng.normal(size=n) > 0).astype(np.float32)
    p = dict(objective="binary", num_leaves=7, use_quantized_grad=True, num_grad_quant_bins=16,
             quant_train_renew_leaf=True, bagging_fraction=fraction, bagging_freq=1, bagging_seed=2)
    ob = O.Booster(O.Dataset(X, y, p), p)
    ob.update()
    ob.update()
    bag = ob.bag()
    gq, hq = ob.quant_packed()
    gs, hs = ob.quant_scales()
    sg, sh, isum = ob.root_sums()
    packed = (gq.astype(np.int64) << 32) | hq.view(np.uint8).astype(np.int64)
    m = len(bag)
    if fraction <= 0.5:
        want_g = want_h = 0.0
        for r in bag:
            want_g += float(gq[r]) * gs
            want_h += float(hq[r]) * hs
        want_i = int(packed[bag].sum())
    else:
        g32, h32 = np.float32(gs), np.float32(hs)
        want_g = want_h = 0.0
        for r in bag:
            want_g += float(np.float32(gq[r]) * g32)
            want_h += float(np.float32(hq[r]) * h32)
        want_i = int(packed[:m].sum())          # by position, not over the bag
        assert want_i != int(packed[bag].sum())
    assert (sg, sh, isum) == (want_g, want_h, want_i)
