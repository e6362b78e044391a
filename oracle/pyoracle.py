"""ctypes binding of the CPU oracle (test infrastructure only).

Loaded by tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` as the checker; the product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

P = C.c_void_p
I = C.c_int
I64 = C.c_int64
D = C.c_double
PF = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
PD = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
PI = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
PU = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
PU8 = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
PU16 = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
PI8 = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")
PI64 = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        sig = {
            "orc_dataset_create": (P, [PF, I64, I, PF, C.c_char_p]),
            "orc_dataset_free": (None, [P]),
            "orc_dataset_num_features": (I, [P]),
            "orc_dataset_num_groups": (I, [P]),
            "orc_dataset_feature": (I, [P, I, C.POINTER(I), C.POINTER(I), C.POINTER(I), C.POINTER(I), C.POINTER(I), PD]),
            "orc_dataset_bins": (None, [P, PU16]),
            "orc_booster_create": (P, [P, C.c_char_p]),
            "orc_booster_free": (None, [P]),
            "orc_booster_train_one_iter": (I, [P]),
            "orc_booster_num_trees": (I, [P]),
            "orc_booster_scores": (None, [P, PD]),
            "orc_booster_gradients": (None, [P, PF, PF]),
            "orc_tree_num_leaves": (I, [P, I]),
            "orc_tree_arrays": (None, [P, I, PI, PU, PD, PI8, PI, PI, PF, PD, PI, PD, PD]),
            "orc_tree_trace": (I, [P, I, PI, PI, PU, PD]),
            "orc_learner_train": (I, [P, PF, PF]),
            "orc_learner_leaf_of_row": (None, [P, PI]),
            "orc_learner_quant_scales": (None, [P, C.POINTER(D), C.POINTER(D)]),
            "orc_learner_quant_packed": (None, [P, PI8, PI8]),
            "orc_booster_predict_raw": (None, [P, PF, I64, I, PD]),
            "orc_hist_f64": (None, [PU8, I, PI, I, PU, I64, PF, PF, PD]),
            "orc_hist_int": (None, [PU8, I, PI, I, PU, I64, PI8, PI8, PI64]),
            "orc_hist_fix": (None, [I, I, D, D, PD]),
            "orc_hist_fix_int": (None, [I, I, I64, PI64]),
            "orc_find_best_threshold": (None, [PD, I, I, I, I, I, PD, PI, D, D, I, D, PD, PI, PI]),
            "orc_find_best_threshold_int": (None, [PI64, I, I, I, I, I, PD, PI, I64, D, D, I, D, PD, PI, PI64, PI]),
            "orc_binary_gradients": (None, [PD, PF, I64, D, PF, PF]),
            "orc_multiclass_gradients": (None, [PD, PF, I64, I, PF, PF]),
            "orc_discretize": (None, [PF, PF, I64, I, I, I, I, PI8, PI8, PD]),
            "orc_quant_tables": (None, [I64, I, PD, PD, I, PI]),
            "orc_to_i8": (I, [D]),
            "orc_mt19937_nth": (C.c_uint32, [C.c_uint32, I]),
            "orc_random_sample": (I, [I, I, I, PI]),
            "orc_find_bin": (I, [PD, I, I64, I, I, I, I, I, PD, C.POINTER(I), C.POINTER(I), C.POINTER(I), C.POINTER(I), PI]),
            "orc_value_to_bin_num": (I, [PD, I, I, D]),
            "orc_auc": (D, [PD, PF, I64]),
            "orc_dataset_set_query": (None, [P, PI, I]),
            "orc_lambdarank_gradients": (None, [PD, PF, PI, I, D, I, I, PF, PF]),
            "orc_ndcg": (D, [PD, PF, PI, I, I]),
            "orc_max_dcg_at_k": (D, [PF, I, I]),
            "orc_binary_logloss": (D, [PD, PF, I64, D]),
            "orc_booster_bag": (I64, [P, P]),
            "orc_booster_predict_window": (None, [P, PF, I64, I, I, I, I, D, PD]),
            "orc_metric_auc": (D, [PD, PF, P, I64]),
            "orc_metric_point": (D, [I, PD, PF, P, I64, I, D, I]),
            "orc_metric_ndcg": (None, [PD, PF, PI, I, PI, I, I, PD]),
            "orc_model_load": (P, [C.c_char_p]),
            "orc_set_num_threads": (None, [I]),
            "orc_learner_root_sums": (None, [P, C.POINTER(D), C.POINTER(D), C.POINTER(I64)]),
            "orc_goss": (I64, [PF, PF, I64, I, I, D, D, I, PU]),
            "orc_block_info_force_size": (None, [I, I64, I64, C.POINTER(I64), C.POINTER(I64)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def params_str(params):
    parts = []
    for k, v in params.items():
        if isinstance(v, bool):
            v = "true" if v else "false"
        elif isinstance(v, (list, tuple)):
            v = ":".join(str(x) for x in v)
        parts.append(f"{k}={v}")
    return " ".join(parts).encode()


class FeatureInfo:
    __slots__ = ("real", "num_bin", "missing_type", "default_bin", "most_freq_bin", "bin_type", "upper")


class Dataset:
    """Oracle dataset: host binning restated from the reference."""

    def __init__(self, X, label, params):
        X = np.ascontiguousarray(X, dtype=np.float32)
        self.n, self.ncol = X.shape
        lab = np.ascontiguousarray(label, dtype=np.float32) if label is not None else np.zeros(self.n, np.float32)
        self.h = lib().orc_dataset_create(X, self.n, self.ncol, lab, params_str(params))
        if not self.h:
            raise ValueError("oracle rejected the parameters")
        self.num_features = lib().orc_dataset_num_features(self.h)
        self.num_groups = lib().orc_dataset_num_groups(self.h)
        self.features = []
        for f in range(self.num_features):
            r, mt, db, mfb, bt = I(), I(), I(), I(), I()
            up = np.zeros(70000, np.float64)
            nb = lib().orc_dataset_feature(self.h, f, C.byref(r), C.byref(mt), C.byref(db), C.byref(mfb), C.byref(bt), up)
            fi = FeatureInfo()
            fi.real, fi.num_bin, fi.missing_type = r.value, nb, mt.value
            fi.default_bin, fi.most_freq_bin, fi.bin_type = db.value, mfb.value, bt.value
            fi.upper = up[:nb].copy()
            self.features.append(fi)

    def set_query_boundaries(self, qb):
        qb = np.ascontiguousarray(qb, np.int32)
        lib().orc_dataset_set_query(self.h, qb, len(qb) - 1)

    def bins(self):
        out = np.zeros((self.n, self.num_features), np.uint16)
        lib().orc_dataset_bins(self.h, out)
        return out

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_dataset_free(self.h)
            self.h = None


class Tree:
    pass


class Booster:
    def __init__(self, ds, params):
        self.ds = ds
        self.params = dict(params)
        self.K = int(params.get("num_class", 1)) if params.get("objective") == "multiclass" else 1
        self.h = lib().orc_booster_create(ds.h, params_str(params))
        if not self.h:
            raise ValueError("oracle rejected the parameters")

    def update(self):
        return lib().orc_booster_train_one_iter(self.h) != 0

    def num_trees(self):
        return lib().orc_booster_num_trees(self.h)

    def scores(self):
        out = np.zeros(self.K * self.ds.n, np.float64)
        lib().orc_booster_scores(self.h, out)
        return out

    def gradients(self):
        g = np.zeros(self.K * self.ds.n, np.float32)
        h = np.zeros(self.K * self.ds.n, np.float32)
        lib().orc_booster_gradients(self.h, g, h)
        return g, h

    def tree(self, t):
        nl = lib().orc_tree_num_leaves(self.h, t)
        ni = max(nl - 1, 1)
        T = Tree()
        T.num_leaves = nl
        T.split_feature = np.zeros(ni, np.int32)
        T.threshold_in_bin = np.zeros(ni, np.uint32)
        T.threshold = np.zeros(ni, np.float64)
        T.decision_type = np.zeros(ni, np.int8)
        T.left_child = np.zeros(ni, np.int32)
        T.right_child = np.zeros(ni, np.int32)
        T.split_gain = np.zeros(ni, np.float32)
        T.leaf_value = np.zeros(nl, np.float64)
        T.leaf_count = np.zeros(nl, np.int32)
        T.leaf_weight = np.zeros(nl, np.float64)
        T.internal_value = np.zeros(ni, np.float64)
        lib().orc_tree_arrays(self.h, t, T.split_feature, T.threshold_in_bin, T.threshold, T.decision_type,
                              T.left_child, T.right_child, T.split_gain, T.leaf_value, T.leaf_count,
                              T.leaf_weight, T.internal_value)
        for a in ("split_feature", "threshold_in_bin", "threshold", "decision_type", "left_child", "right_child",
                  "split_gain", "internal_value"):
            setattr(T, a, getattr(T, a)[: nl - 1])
        return T

    def trace(self, t):
        n = 4096
        leaf, feat = np.zeros(n, np.int32), np.zeros(n, np.int32)
        thr, gain = np.zeros(n, np.uint32), np.zeros(n, np.float64)
        k = lib().orc_tree_trace(self.h, t, leaf, feat, thr, gain)
        return leaf[:k], feat[:k], thr[:k], gain[:k]

    def train_tree(self, g, h):
        """One tree from caller gradients (no score update)."""
        return lib().orc_learner_train(self.h, np.ascontiguousarray(g, np.float32), np.ascontiguousarray(h, np.float32))

    def leaf_of_row(self):
        out = np.zeros(self.ds.n, np.int32)
        lib().orc_learner_leaf_of_row(self.h, out)
        return out

    def quant_scales(self):
        a, b = D(), D()
        lib().orc_learner_quant_scales(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def root_sums(self):
        """The last tree's root sums: (gradient, hessian, packed integer)."""
        a, b, c = D(), D(), I64()
        lib().orc_learner_root_sums(self.h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def quant_packed(self):
        gq = np.zeros(self.ds.n, np.int8)
        hq = np.zeros(self.ds.n, np.int8)
        lib().orc_learner_quant_packed(self.h, gq, hq)
        return gq, hq

    def predict_window(self, X, start_iteration=0, num_iteration=-1, freq=0, margin=0.0):
        X = np.ascontiguousarray(X, np.float32)
        out = np.zeros(X.shape[0] * self.K, np.float64)
        lib().orc_booster_predict_window(self.h, X, X.shape[0], X.shape[1], start_iteration, num_iteration, freq,
                                         margin, out)
        return out.reshape(X.shape[0], self.K)

    def bag(self):
        """The current bag (in-bag rows ascending), or None without bagging."""
        k = lib().orc_booster_bag(self.h, None)
        if k < 0:
            return None
        out = np.zeros(max(k, 1), np.uint32)
        lib().orc_booster_bag(self.h, out.ctypes.data)
        return out[:k]

    def predict_raw(self, X):
        X = np.ascontiguousarray(X, np.float32)
        out = np.zeros(X.shape[0] * self.K, np.float64)
        lib().orc_booster_predict_raw(self.h, X, X.shape[0], X.shape[1], out)
        return out.reshape(X.shape[0], self.K)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_booster_free(self.h)
            self.h = None


class LoadedModel(Booster):
    """A prediction-only oracle booster from model text (io.rs:156-343):
    predict_raw and predict_window of Booster on the same trees."""

    def __init__(self, text, K=None):
        self.ds = None
        self.h = lib().orc_model_load(text.encode())
        if not self.h:
            raise ValueError("oracle could not read the model text")
        head = dict(l.split("=", 1) for l in text.split("Tree=")[0].splitlines() if "=" in l)
        self.K = int(head.get("num_tree_per_iteration", head.get("num_class", "1")))


def block_info_force_size(num_threads, n, min_size):
    a, b = I64(), I64()
    lib().orc_block_info_force_size(num_threads, n, min_size, C.byref(a), C.byref(b))
    return a.value, b.value


def set_num_threads(n):
    """Host threads of the oracle's loops (trees and scores do not depend on it)."""
    lib().orc_set_num_threads(int(n))


def goss(g, h, K, num_threads, top_rate, other_rate, seed):
    """One GOSS draw on gradients [K][n] (copies): (in-bag rows, rescaled g, rescaled h)."""
    g = np.ascontiguousarray(g, np.float32).copy()
    h = np.ascontiguousarray(h, np.float32).copy()
    n = g.size // K
    bag = np.zeros(max(n, 1), np.uint32)
    m = lib().orc_goss(g, h, n, K, num_threads, top_rate, other_rate, seed, bag)
    return bag[:m], g, h


def hist_f64(bins_u8, mfb, B, idx, g, h):
    n, F = bins_u8.shape
    out = np.zeros(F * B * 2, np.float64)
    lib().orc_hist_f64(np.ascontiguousarray(bins_u8, np.uint8), F, np.ascontiguousarray(mfb, np.int32), B,
                       np.ascontiguousarray(idx, np.uint32), len(idx), np.ascontiguousarray(g, np.float32),
                       np.ascontiguousarray(h, np.float32), out)
    return out.reshape(F, B, 2)


def hist_int(bins_u8, mfb, B, idx, gq, hq):
    n, F = bins_u8.shape
    out = np.zeros(F * B, np.int64)
    lib().orc_hist_int(np.ascontiguousarray(bins_u8, np.uint8), F, np.ascontiguousarray(mfb, np.int32), B,
                       np.ascontiguousarray(idx, np.uint32), len(idx), np.ascontiguousarray(gq, np.int8),
                       np.ascontiguousarray(hq, np.int8), out)
    return out.reshape(F, B)


def split_params_arrays(p):
    dp = np.array([p.get("lambda_l1", 0.0), p.get("lambda_l2", 0.0), p.get("max_delta_step", 0.0),
                   p.get("path_smooth", 0.0), p.get("min_sum_hessian_in_leaf", 1e-3), p.get("min_gain_to_split", 0.0),
                   p.get("cat_l2", 10.0), p.get("cat_smooth", 10.0)], np.float64)
    ip = np.array([p.get("min_data_in_leaf", 20), p.get("max_cat_to_onehot", 4), p.get("max_cat_threshold", 32),
                   p.get("min_data_per_group", 100)], np.int32)
    return dp, ip


def hist_fix(hist_bins, mfb, sg, sh):
    """histogram.rs:42-56 on a bin-indexed [B][2] histogram; returns the fixed copy."""
    hb = np.ascontiguousarray(hist_bins, np.float64).reshape(-1).copy()
    lib().orc_hist_fix(len(hb) // 2, mfb, sg, sh, hb)
    return hb


def hist_fix_int(hist_bins, mfb, total):
    """histogram.rs:67-78 on a bin-indexed packed int64 histogram; returns the fixed copy."""
    hb = np.ascontiguousarray(hist_bins, np.int64).copy()
    lib().orc_hist_fix_int(len(hb), mfb, int(total), hb)
    return hb


def find_best_threshold(hist_bins, num_bin, missing_type, default_bin, mfb, bin_type, p, sg, sh, nd, parent_output=0.0):
    """hist_bins: [B][2] bin-indexed (modified in place by the mfb fix)."""
    dp, ip = split_params_arrays(p)
    hb = np.ascontiguousarray(hist_bins, np.float64).reshape(-1).copy()
    od = np.zeros(7, np.float64)
    oi = np.zeros(5, np.int32)
    oc = np.zeros(num_bin + 1, np.int32)
    lib().orc_find_best_threshold(hb, num_bin, missing_type, default_bin, mfb, bin_type, dp, ip, sg, sh, nd,
                                  parent_output, od, oi, oc)
    return dict(gain=od[0], left_output=od[1], right_output=od[2], left_sum_gradient=od[3], left_sum_hessian=od[4],
                right_sum_gradient=od[5], right_sum_hessian=od[6], threshold=int(oi[0]), default_left=bool(oi[1]),
                left_count=int(oi[2]), right_count=int(oi[3]), splittable=bool(oi[4]),
                cat_threshold=[int(v) for v in oc[1:1 + oc[0]]])


def find_best_threshold_int(hist_bins, num_bin, missing_type, default_bin, mfb, p, int_sum, gscale, hscale, nd,
                            parent_output=0.0, bin_type=0):
    dp, ip = split_params_arrays(p)
    hb = np.ascontiguousarray(hist_bins, np.int64).reshape(-1).copy()
    od = np.zeros(7, np.float64)
    oi = np.zeros(5, np.int32)
    og = np.zeros(2, np.int64)
    oc = np.zeros(num_bin + 1, np.int32)
    lib().orc_find_best_threshold_int(hb, num_bin, missing_type, default_bin, mfb, bin_type, dp, ip, int_sum, gscale,
                                      hscale, nd, parent_output, od, oi, og, oc)
    return dict(gain=od[0], left_output=od[1], right_output=od[2], left_sum_gradient=od[3], left_sum_hessian=od[4],
                right_sum_gradient=od[5], right_sum_hessian=od[6], threshold=int(oi[0]), default_left=bool(oi[1]),
                left_count=int(oi[2]), right_count=int(oi[3]), splittable=bool(oi[4]), left_sum_gh=int(og[0]),
                right_sum_gh=int(og[1]), cat_threshold=[int(v) for v in oc[1:1 + oc[0]]])


def binary_gradients(scores, label, sigmoid=1.0):
    n = len(label)
    g = np.zeros(n, np.float32)
    h = np.zeros(n, np.float32)
    lib().orc_binary_gradients(np.ascontiguousarray(scores, np.float64), np.ascontiguousarray(label, np.float32), n,
                               sigmoid, g, h)
    return g, h


def multiclass_gradients(scores, label, K):
    n = len(label)
    g = np.zeros(n * K, np.float32)
    h = np.zeros(n * K, np.float32)
    lib().orc_multiclass_gradients(np.ascontiguousarray(scores, np.float64), np.ascontiguousarray(label, np.float32),
                                   n, K, g, h)
    return g, h


def lambdarank_gradients(scores, label, qb, sigmoid=1.0, norm=True, trunc=30):
    n = len(label)
    g = np.zeros(n, np.float32)
    h = np.zeros(n, np.float32)
    qb = np.ascontiguousarray(qb, np.int32)
    lib().orc_lambdarank_gradients(np.ascontiguousarray(scores, np.float64), np.ascontiguousarray(label, np.float32),
                                   qb, len(qb) - 1, sigmoid, int(norm), trunc, g, h)
    return g, h


def ndcg(scores, label, qb, k=10):
    qb = np.ascontiguousarray(qb, np.int32)
    return lib().orc_ndcg(np.ascontiguousarray(scores, np.float64), np.ascontiguousarray(label, np.float32), qb,
                          len(qb) - 1, k)


def max_dcg_at_k(label, k):
    lab = np.ascontiguousarray(label, np.float32)
    return lib().orc_max_dcg_at_k(lab, len(lab), k)


def discretize(g, h, bins, seed=0, stochastic=True, num_calls=1):
    n = len(g)
    gq = np.zeros(n, np.int8)
    hq = np.zeros(n, np.int8)
    sc = np.zeros(2, np.float64)
    lib().orc_discretize(np.ascontiguousarray(g, np.float32), np.ascontiguousarray(h, np.float32), n, bins, seed,
                         int(stochastic), num_calls, gq, hq, sc)
    return gq, hq, sc[0], sc[1]


def quant_tables(n, seed, count):
    grv = np.zeros(n, np.float64)
    hrv = np.zeros(n, np.float64)
    st = np.zeros(max(count, 1), np.int32)
    lib().orc_quant_tables(n, seed, grv, hrv, count, st)
    return grv, hrv, st[:count]


def auc(score, label):
    return lib().orc_auc(np.ascontiguousarray(score, np.float64), np.ascontiguousarray(label, np.float32), len(label))


def _f32_or_none(w):
    return None if w is None else np.ascontiguousarray(w, np.float32)


def metric_auc(score, label, weight=None):
    """metric.rs:701-729 (weighted when weight is given)."""
    w = _f32_or_none(weight)
    return lib().orc_metric_auc(np.ascontiguousarray(score, np.float64), np.ascontiguousarray(label, np.float32),
                                None if w is None else w.ctypes.data, len(label))


POINT_KINDS = {"binary_logloss": 0, "binary_error": 1, "multi_logloss": 2, "multi_error": 3}


def metric_point(kind, score, label, weight=None, K=1, sigmoid=1.0, top_k=1):
    """metric.rs:602-653; score class-major [K][n]."""
    w = _f32_or_none(weight)
    return lib().orc_metric_point(POINT_KINDS[kind], np.ascontiguousarray(score, np.float64).ravel(),
                                  np.ascontiguousarray(label, np.float32),
                                  None if w is None else w.ctypes.data, len(label), K,
                                  sigmoid, top_k)


def metric_ndcg(score, label, qb, eval_at=(1, 2, 3, 4, 5), threads=1):
    """metric.rs:551-600: NDCG at each cut-off."""
    at = np.ascontiguousarray(eval_at, np.int32)
    qb = np.ascontiguousarray(qb, np.int32)
    out = np.zeros(len(at), np.float64)
    lib().orc_metric_ndcg(np.ascontiguousarray(score, np.float64), np.ascontiguousarray(label, np.float32), qb,
                          len(qb) - 1, at, len(at), threads, out)
    return out


def binary_logloss(raw, label, sigmoid=1.0):
    return lib().orc_binary_logloss(np.ascontiguousarray(raw, np.float64), np.ascontiguousarray(label, np.float32),
                                    len(label), sigmoid)


def find_bin(vals, total, bin_type=0, max_bin=255, min_data_in_bin=3, min_split_data=0, pre_filter=False):
    v = np.ascontiguousarray(vals, np.float64)
    up = np.zeros(max(max_bin + 2, 300), np.float64)
    b2c = np.zeros(max(max_bin + 2, 300), np.int32)
    mt, db, mfb, tr = I(), I(), I(), I()
    nb = lib().orc_find_bin(v, len(v), total, bin_type, max_bin, min_data_in_bin, min_split_data, int(pre_filter), up,
                            C.byref(mt), C.byref(db), C.byref(mfb), C.byref(tr), b2c)
    return dict(num_bin=nb, upper=up[:nb].copy(), bin_2_categorical=b2c[:nb].copy(), missing_type=mt.value,
                default_bin=db.value, most_freq_bin=mfb.value, is_trivial=bool(tr.value))
