"""MI355X GBDT engine: the lightgbm-rust training/prediction hot path on HIP.

Host-side mirror of the reference's Python API (python/lightgbm_rust/:
``Dataset`` basic.py:710, ``Booster`` basic.py:1364, ``train`` engine.py:50)
for the hot path, over the C ABI of liblgbm_hip.so. The only API change is
the device option: ``device_type`` (alias ``device``) = ``"gpu"`` / ``"hip"``
selects this engine (config.rs:787-793 gates it in the reference).

Binning (the BinMapper) runs on the host; the packed bin matrix is uploaded
to HBM once; gradients, histograms, threshold search, partition, score
update and prediction run as HIP kernels on gfx950.
"""
import collections
import ctypes as C

import numpy as np

from . import _capi
from ._capi import LightGBMError, check, lib, ptr

__all__ = ["Dataset", "Booster", "train", "LightGBMError", "DEFAULT_PARAMS", "EvalResult", "CallbackEnv"]

# lgbm-core defaults of the parameters this path reads (config.rs:508-620).
DEFAULT_PARAMS = dict(
    objective="regression", num_iterations=100, learning_rate=0.1, num_leaves=31, max_depth=-1,
    min_data_in_leaf=20, min_sum_hessian_in_leaf=1e-3, lambda_l1=0.0, lambda_l2=0.0, max_delta_step=0.0,
    min_gain_to_split=0.0, path_smooth=0.0, max_bin=255, min_data_in_bin=3, bin_construct_sample_cnt=200000,
    data_random_seed=1, seed=0, feature_pre_filter=True, use_missing=True, zero_as_missing=False, sigmoid=1.0,
    boost_from_average=True, num_class=1, use_quantized_grad=False, num_grad_quant_bins=4,
    stochastic_rounding=True, quant_train_renew_leaf=False, max_cat_to_onehot=4, max_cat_threshold=32,
    cat_l2=10.0, cat_smooth=10.0, min_data_per_group=100, lambdarank_truncation_level=30, lambdarank_norm=True,
    num_threads=0, device_type="gpu", gpu_device_id=0, use_graph=True, quant_table_threads=None,
    metric="", eval_at=(1, 2, 3, 4, 5), multi_error_top_k=1, label_gain=None, early_stopping_round=0,
    first_metric_only=False, early_stopping_min_delta=0.0, device_binning=False, bagging_fraction=1.0,
    bagging_freq=0, bagging_seed=3, pos_bagging_fraction=1.0, neg_bagging_fraction=1.0,
    data_sample_strategy="bagging", top_rate=0.2, other_rate=0.1, autotune=False, autotune_trees=4, bagging_by_query=False,
    batch_trees=True, pack_u4=True, enable_bundle=True,
)

ALIASES = {
    "device": "device_type", "num_boost_round": "num_iterations", "n_estimators": "num_iterations",
    "num_iteration": "num_iterations", "num_tree": "num_iterations", "num_trees": "num_iterations",
    "eta": "learning_rate", "shrinkage_rate": "learning_rate", "num_leaf": "num_leaves", "max_leaves": "num_leaves",
    "min_data_per_leaf": "min_data_in_leaf", "min_data": "min_data_in_leaf", "min_child_samples": "min_data_in_leaf",
    "min_sum_hessian_per_leaf": "min_sum_hessian_in_leaf", "min_child_weight": "min_sum_hessian_in_leaf",
    "reg_alpha": "lambda_l1", "l1_regularization": "lambda_l1", "reg_lambda": "lambda_l2", "lambda": "lambda_l2",
    "l2_regularization": "lambda_l2", "min_split_gain": "min_gain_to_split", "max_bins": "max_bin",
    "num_classes": "num_class", "application": "objective", "objective_type": "objective", "app": "objective",
    "random_seed": "seed", "random_state": "seed", "data_seed": "data_random_seed", "subsample_for_bin":
    "bin_construct_sample_cnt", "num_thread": "num_threads", "nthread": "num_threads", "n_jobs": "num_threads",
    "cat_feature": "categorical_feature", "categorical_column": "categorical_feature",
    "metrics": "metric", "metric_types": "metric", "ndcg_eval_at": "eval_at", "ndcg_at": "eval_at",
    "map_eval_at": "eval_at", "map_at": "eval_at", "early_stopping_rounds": "early_stopping_round",
    "early_stopping": "early_stopping_round", "n_iter_no_change": "early_stopping_round",
    "sub_row": "bagging_fraction", "subsample": "bagging_fraction", "bagging": "bagging_fraction",
    "subsample_freq": "bagging_freq", "bagging_fraction_seed": "bagging_seed",
    "pos_sub_row": "pos_bagging_fraction", "pos_subsample": "pos_bagging_fraction", "pos_bagging": "pos_bagging_fraction",
    "neg_sub_row": "neg_bagging_fraction", "neg_subsample": "neg_bagging_fraction", "neg_bagging": "neg_bagging_fraction",
}

# metric names and aliases on this path (config.rs ParseMetrics, metric.rs:44-76)
METRIC_ALIASES = {
    "binary_logloss": "binary_logloss", "binary": "binary_logloss", "binary_error": "binary_error",
    "auc": "auc", "multi_logloss": "multi_logloss", "multiclass": "multi_logloss", "softmax": "multi_logloss",
    "multiclassova": "multi_logloss", "multiclass_ova": "multi_logloss", "ova": "multi_logloss",
    "ovr": "multi_logloss", "multi_error": "multi_error", "ndcg": "ndcg", "lambdarank": "ndcg",
    "rank_xendcg": "ndcg", "xendcg": "ndcg", "xe_ndcg": "ndcg", "xe_ndcg_mart": "ndcg", "xendcg_mart": "ndcg",
}
DEFAULT_METRIC = {"binary": "binary_logloss", "multiclass": "multi_logloss", "softmax": "multi_logloss",
                  "lambdarank": "ndcg"}
HIGHER_BETTER = {"auc", "ndcg"}

EvalResult = collections.namedtuple("EvalResult", ["dataset_name", "metric_name", "metric_value", "maximize"])
CallbackEnv = collections.namedtuple(
    "CallbackEnv", ["model", "params", "iteration", "begin_iteration", "end_iteration", "evaluation_result_list"])


def _metrics(p):
    m = p.get("metric")
    names = [x.strip() for x in m.split(",")] if isinstance(m, str) else list(m or [])
    names = [x for x in names if x]
    if not names:
        names = [DEFAULT_METRIC[str(p["objective"])]]
    out = []
    for x in names:
        if x in ("None", "none", "null", "na", "custom"):
            return []
        if x not in METRIC_ALIASES:
            raise LightGBMError(f"metric {x!r} is not evaluated on the GPU path")
        if METRIC_ALIASES[x] not in out:
            out.append(METRIC_ALIASES[x])
    return out

OBJECTIVES = {"binary": 0, "multiclass": 1, "softmax": 1, "lambdarank": 2}
DEVICES = {"gpu", "hip", "cuda"}


def _lcg_short(x):
    """random.rs:19-33: the next state and next_short(0, 32767) of the upstream LCG."""
    x = (x * 214013 + 2531011) & 0xFFFFFFFF
    return x, ((x >> 16) & 0x7FFF) % 32767


def _params(params):
    out = dict(DEFAULT_PARAMS)
    given = {ALIASES.get(k, k): v for k, v in (params or {}).items()}
    if "seed" in given:
        # config.rs:795-806: an explicit seed derives data_random_seed and
        # bagging_seed (explicit values of those still win)
        x = int(given["seed"]) & 0xFFFFFFFF
        x, out["data_random_seed"] = _lcg_short(x)
        x, out["bagging_seed"] = _lcg_short(x)
    for k, v in given.items():
        out[k] = v
    if str(out["device_type"]).lower() not in DEVICES:
        raise LightGBMError("device_type must be gpu/hip for this engine (the reference's CPU engine handles cpu)")
    return out


def _b(v):
    if isinstance(v, str):
        return v.lower() in ("true", "1", "+", "yes")
    return bool(v)


class Dataset:
    """Binned training data (basic.py:710). Host binning; bins uploaded lazily by Booster."""

    def __init__(self, data, label=None, weight=None, group=None, params=None, categorical_feature=None,
                 reference=None, free_raw_data=True):
        self.params = _params(params)
        X = np.ascontiguousarray(data, dtype=np.float32)
        if X.ndim != 2:
            raise LightGBMError("data must be 2-D")
        self.num_data, self.num_total_features = X.shape
        self.label = None if label is None else np.ascontiguousarray(label, dtype=np.float32)
        if self.label is not None and len(self.label) != self.num_data:
            raise LightGBMError(f"Length of label ({len(self.label)}) is not same with #data ({self.num_data})")
        self.weight = None if weight is None else np.ascontiguousarray(weight, dtype=np.float32)
        self.group = None if group is None else np.asarray(group, dtype=np.int32)
        cats = categorical_feature if categorical_feature is not None else self.params.get("categorical_feature")
        if isinstance(cats, str):
            cats = [int(c) for c in cats.split(",") if c.strip()] if cats else []
        self.categorical = np.ascontiguousarray(cats or [], dtype=np.int32)
        self.reference = reference
        self._raw = None
        self._bins = None
        if reference is not None:
            self._handle = reference._handle
            self._owns = False
            self._bins = np.zeros((self.num_data, reference.num_features), np.uint8)
            check(lib().lgbmhip_bins_apply(self._handle, ptr(X), self.num_data, self.num_total_features,
                                           ptr(self._bins)))
            self.metas, self.upper = reference.metas, reference.upper
            self.num_features = reference.num_features
            # a dataset binned with another's mappers keeps one group per
            # feature (dataset.rs:628-635, Layout::Valid)
            self.num_feature_groups = self.num_features
            return
        p = self.params
        bp = _capi.BinParams(int(p["max_bin"]), int(p["min_data_in_bin"]), int(p["min_data_in_leaf"]),
                             int(p["bin_construct_sample_cnt"]), int(p["data_random_seed"]),
                             int(_b(p["feature_pre_filter"])), int(_b(p["use_missing"])),
                             int(_b(p["zero_as_missing"])), int(p["num_threads"]))
        h = C.c_void_p()
        # device_binning: only the sample-based mappers on the host; the
        # Booster bins the rows on the GPU (lgbmhip_upload_raw)
        device = _b(p["device_binning"])
        find = lib().lgbmhip_find_mappers if device else lib().lgbmhip_bin_dense
        check(find(ptr(X), self.num_data, self.num_total_features,
                   ptr(self.categorical) if len(self.categorical) else None, len(self.categorical), C.byref(bp),
                   C.byref(h)))
        self._handle = h
        self._owns = True
        self.num_features = lib().lgbmhip_bins_num_features(h)
        # the reference's feature-group count (exclusive feature bundling,
        # feature_groups.rs:169-211); one group per feature without bundling
        self.num_feature_groups = (lib().lgbmhip_bins_num_groups(h) if _b(p["enable_bundle"])
                                   else self.num_features)
        self.metas = (_capi.FeatureMeta * max(self.num_features, 1))()
        self.upper = np.zeros(max(lib().lgbmhip_bins_num_upper(h), 1), np.float64)
        check(lib().lgbmhip_bins_metas(h, C.cast(self.metas, C.c_void_p), ptr(self.upper)))
        if device:
            self._raw = X
            return
        self._bins = np.zeros((self.num_data, self.num_features), np.uint8)
        check(lib().lgbmhip_bins_matrix(h, ptr(self._bins)))

    @property
    def bins(self):
        """The n x F u8 bin matrix (binned on the host on first use when the
        rows are binned on the device for training)."""
        if self._bins is None:
            self._bins = np.zeros((self.num_data, self.num_features), np.uint8)
            check(lib().lgbmhip_bins_apply(self._handle, ptr(self._raw), self.num_data, self.num_total_features,
                                           ptr(self._bins)))
        return self._bins

    def feature_infos(self):
        """Dataset::feature_infos: '[min:max]', categories 'a:b', or 'none' per raw column."""
        n = C.c_int64(0)
        check(lib().lgbmhip_bins_feature_infos(self._handle, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value + 1)
        check(lib().lgbmhip_bins_feature_infos(self._handle, buf, n.value + 1, C.byref(n)))
        return buf.value.decode()

    def feature_meta(self, f):
        m = self.metas[f]
        return dict(real_index=m.real_index, num_bin=m.num_bin, missing_type=m.missing_type,
                    default_bin=m.default_bin, most_freq_bin=m.most_freq_bin, bin_type=m.bin_type)

    def construct(self):
        return self

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h and getattr(self, "_owns", False) and _capi._lib is not None:
            _capi._lib.lgbmhip_bins_free(h)
            self._handle = None


class Tree:
    """Tree arrays (tree.rs:120-159)."""


def _batch_trees(v):
    """Class trees per shared growth launch: True = the engine's choice,
    False or 0 = one tree per launch (growth lanes), n >= 2 = batches of n."""
    if isinstance(v, str):
        v = {"true": True, "false": False}.get(v.lower(), v)
    if isinstance(v, bool):
        return 1 if v else 0
    return max(0, int(v))


def _quant_table_threads(p):
    """The thread count the quantizer's random tables are drawn for
    (learner/quant.rs:45-79: one generator pair per thread block): the given
    quant_table_threads, else num_threads when it is set, else 1 (the
    reference's default of all host threads has no fixed value here)."""
    if p.get("quant_table_threads") is not None:
        return int(p["quant_table_threads"])
    return int(p["num_threads"]) if int(p["num_threads"]) > 0 else 1


class Booster:
    """Training booster on the GPU engine (basic.py:1364 / py/lib.rs:435 RsBooster)."""

    def __init__(self, params=None, train_set=None, model_file=None, model_str=None):
        if train_set is None:
            if model_file is None and model_str is None:
                raise LightGBMError("Booster needs a train_set, a model_file or a model_str")
            self._load(params, model_file, model_str)
            return
        self.params = _params({**(train_set.params or {}), **(params or {})})
        p = self.params
        obj = str(p["objective"])
        if obj not in OBJECTIVES:
            raise LightGBMError(f"objective {obj!r} is not on the GPU path (binary, multiclass, lambdarank)")
        if train_set.label is None:
            raise LightGBMError("train_set has no label")
        self.train_set = train_set
        self.K = int(p["num_class"]) if OBJECTIVES[obj] == 1 else 1
        h = C.c_void_p()
        check(lib().lgbmhip_create(int(p["gpu_device_id"]), C.byref(h)))
        self._ctx = h
        # features of at most 16 bins stored as 4-bit fields in HBM (bin.rs:92-121 Dense4)
        check(lib().lgbmhip_set_bin_packing(h, int(_b(p["pack_u4"]))), h)
        ds = train_set
        if ds._raw is not None:
            check(lib().lgbmhip_upload_raw(h, ds._handle, ptr(ds._raw), ds.num_data, ds.num_total_features, 0), h)
        else:
            check(lib().lgbmhip_upload_bins(h, C.cast(ds.metas, C.c_void_p), ds.num_features, ptr(ds.upper),
                                            ptr(ds.bins), ds.num_data, ds.num_total_features), h)
        # the bagging subset-mode rule reads the reference's feature-group count
        check(lib().lgbmhip_set_feature_groups(h, int(ds.num_feature_groups)), h)
        qb = None
        nq = 0
        if ds.group is not None:
            qb = np.ascontiguousarray(np.concatenate([[0], np.cumsum(ds.group)]), dtype=np.int32)
            nq = len(ds.group)
        check(lib().lgbmhip_set_metadata(h, ptr(ds.label), ptr(ds.weight), ptr(qb), nq), h)
        sp = _capi.SplitParams(float(p["lambda_l1"]), float(p["lambda_l2"]), float(p["max_delta_step"]),
                               float(p["path_smooth"]), float(p["min_sum_hessian_in_leaf"]),
                               float(p["min_gain_to_split"]), float(p["cat_l2"]), float(p["cat_smooth"]),
                               int(p["min_data_in_leaf"]), int(p["max_cat_to_onehot"]), int(p["max_cat_threshold"]),
                               int(p["min_data_per_group"]), int(p["num_leaves"]), int(p["max_depth"]))
        check(lib().lgbmhip_set_split_params(h, C.byref(sp)), h)
        to = _capi.TrainOpts(OBJECTIVES[obj], self.K, float(p["sigmoid"]), float(p["learning_rate"]),
                             int(_b(p["boost_from_average"])), int(_b(p["use_quantized_grad"])),
                             int(p["num_grad_quant_bins"]), int(_b(p["stochastic_rounding"])),
                             int(_b(p["quant_train_renew_leaf"])), int(p["seed"]), _quant_table_threads(p),
                             int(p["lambdarank_truncation_level"]), int(_b(p["lambdarank_norm"])),
                             int(_b(p["use_graph"])), float(p["bagging_fraction"]), int(p["bagging_freq"]),
                             int(p["bagging_seed"]), float(p["pos_bagging_fraction"]),
                             float(p["neg_bagging_fraction"]),
                             1 if str(p["data_sample_strategy"]) == "goss" or str(p.get("boosting")) == "goss" else 0,
                             int(p["num_threads"]) if int(p["num_threads"]) > 0 else 4096,
                             float(p["top_rate"]), float(p["other_rate"]), int(_b(p["bagging_by_query"])),
                             _batch_trees(p["batch_trees"]))
        check(lib().lgbmhip_set_train_opts(h, C.byref(to)), h)
        self.tuning = None
        if _b(p["autotune"]):
            # launch shapes timed on this workload's first trees (cached per process and shape)
            out = np.zeros(5, np.int32)
            check(lib().lgbmhip_autotune(h, int(p["autotune_trees"]), ptr(out)), h)
            self.tuning = dict(zip(("part_grid", "hist_chunks", "hist_min_rows", "hist_lds", "candidates_timed"),
                                   out.tolist()))
        self._finished = False
        self._metrics = _metrics(p)
        self._valid = []          # (name, index, Dataset)
        self._train_name = "training"
        self.best_iteration = 0
        self.best_score = {}

    def _load(self, params, model_file, model_str):
        """A prediction-only booster from model text (GBDT::LoadModelFromString)."""
        if model_str is None:
            with open(model_file, "r") as f:
                model_str = f.read()
        self.params = _params(params)
        h = C.c_void_p()
        check(lib().lgbmhip_create(int(self.params["gpu_device_id"]), C.byref(h)))
        self._ctx = h
        check(lib().lgbmhip_load_model(h, model_str.encode()), h)
        head = dict(l.split("=", 1) for l in model_str.split("end of trees")[0].splitlines() if "=" in l and
                    not l.startswith("Tree="))
        self.K = int(head["num_tree_per_iteration"])
        obj = head.get("objective", "").split(" ")
        self.params["objective"] = obj[0] if obj[0] else "binary"
        for tok in obj[1:]:
            if tok.startswith("sigmoid:"):
                self.params["sigmoid"] = float(tok.split(":", 1)[1])
        self.train_set = None
        self._valid, self._metrics = [], []
        self.best_iteration = 0
        self.best_score = {}

    def model_to_string(self, num_iteration=None, parameters=None):
        """Model text (GBDT::SaveModelToString). `parameters`: the parameters
        block's `[name: value]` lines (the config dump the host owns); a model
        loaded from text keeps its own block."""
        ni = -1 if num_iteration is None else int(num_iteration)
        infos = self.train_set.feature_infos().encode() if self.train_set is not None else None
        par = parameters.encode() if parameters is not None else None
        n = C.c_int64(0)
        check(lib().lgbmhip_save_model(self._ctx, ni, None, infos, par, None, 0, C.byref(n)), self._ctx)
        buf = C.create_string_buffer(n.value + 1)
        check(lib().lgbmhip_save_model(self._ctx, ni, None, infos, par, buf, n.value + 1, C.byref(n)), self._ctx)
        return buf.value.decode()

    def save_model(self, filename, num_iteration=None):
        with open(filename, "w") as f:
            f.write(self.model_to_string(num_iteration))
        return self

    # --- training (basic.py:1519 update -> py/lib.rs:479-482 -> Gbdt::train_one_iter)
    def update(self):
        """One boosting iteration; returns True when training cannot continue."""
        fin = C.c_int32(0)
        check(lib().lgbmhip_train_iteration(self._ctx, C.byref(fin)), self._ctx)
        self._finished = bool(fin.value)
        return self._finished

    def train_tree(self, grad, hess):
        """One tree from caller gradients (SerialTreeLearner::train), no score update."""
        g = np.ascontiguousarray(grad, np.float32)
        h = np.ascontiguousarray(hess, np.float32)
        check(lib().lgbmhip_train_tree(self._ctx, ptr(g), ptr(h)), self._ctx)

    def num_trees(self):
        return lib().lgbmhip_num_trees(self._ctx)

    def num_model_per_iteration(self):
        return self.K

    def rollback_one_iter(self):
        """Drops the last iteration's trees from the model and the scores (Gbdt::rollback_one_iter)."""
        check(lib().lgbmhip_rollback_one_iter(self._ctx), self._ctx)
        return self

    def feature_importance(self, importance_type="split", iteration=None):
        """Split counts or total gains per raw column (boosting.rs:1337-1360)."""
        if importance_type not in ("split", "gain"):
            raise LightGBMError("Unknown importance type: only support split=0 and gain=1")
        ncol = self.train_set.num_total_features if self.train_set is not None else None
        n_models = self.num_trees()
        if iteration is not None and int(iteration) > 0:
            n_models = min(n_models, int(iteration) * self.K)
        imp = np.zeros(ncol if ncol is not None else 0, np.float64)
        for t in range(n_models):
            tr = self.tree(t)
            for f, g in zip(tr.split_feature, tr.split_gain):
                if g > 0:
                    if f >= len(imp):
                        imp = np.pad(imp, (0, f + 1 - len(imp)))
                    imp[f] += 1.0 if importance_type == "split" else float(g)
        return imp

    def current_iteration(self):
        return self.num_trees() // self.K

    def tree(self, t):
        nl = lib().lgbmhip_tree_num_leaves(self._ctx, t)
        if nl < 0:
            raise LightGBMError("tree index out of range")
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
        check(lib().lgbmhip_tree_arrays(self._ctx, t, ptr(T.split_feature), ptr(T.threshold_in_bin),
                                        ptr(T.threshold), ptr(T.decision_type), ptr(T.left_child),
                                        ptr(T.right_child), ptr(T.split_gain), ptr(T.leaf_value),
                                        ptr(T.leaf_count), ptr(T.leaf_weight), ptr(T.internal_value)), self._ctx)
        for a in ("split_feature", "threshold_in_bin", "threshold", "decision_type", "left_child", "right_child",
                  "split_gain", "internal_value"):
            setattr(T, a, getattr(T, a)[: nl - 1])
        return T

    def train_scores(self):
        out = np.zeros(self.K * self.train_set.num_data, np.float64)
        check(lib().lgbmhip_copy_scores(self._ctx, ptr(out)), self._ctx)
        return out

    def last_gradients(self):
        n = self.K * self.train_set.num_data
        g = np.zeros(n, np.float32)
        h = np.zeros(n, np.float32)
        check(lib().lgbmhip_copy_gradients(self._ctx, ptr(g), ptr(h)), self._ctx)
        return g, h

    def leaf_of_row(self):
        out = np.zeros(self.train_set.num_data, np.int32)
        check(lib().lgbmhip_leaf_of_row(self._ctx, ptr(out)), self._ctx)
        return out

    # --- validation and metrics (basic.py add_valid / eval_train / eval_valid -> boosting.rs:378-411, 1242-1280)
    def add_valid(self, data, name):
        """Attach a validation Dataset built with reference=train_set."""
        if data.reference is not self.train_set and data is not self.train_set:
            raise LightGBMError("validation data must be constructed with reference to the training data")
        if data.label is None:
            raise LightGBMError("validation data has no label")
        qb, nq = None, 0
        if data.group is not None:
            qb = np.ascontiguousarray(np.concatenate([[0], np.cumsum(data.group)]), dtype=np.int32)
            nq = len(data.group)
        idx = C.c_int32(0)
        check(lib().lgbmhip_add_valid(self._ctx, ptr(data.bins), data.num_data, ptr(data.label), ptr(data.weight),
                                      ptr(qb), nq, None, C.byref(idx)), self._ctx)
        self._valid.append((name, idx.value, data))
        return self

    def set_train_data_name(self, name):
        self._train_name = name
        return self

    def valid_scores(self, i):
        """Raw scores of validation set i (0-based), class-major."""
        name, idx, data = self._valid[i]
        out = np.zeros(self.K * data.num_data, np.float64)
        check(lib().lgbmhip_copy_valid_scores(self._ctx, idx, ptr(out)), self._ctx)
        return out

    def _eval(self, data_index, data_name):
        p = self.params
        res = []
        for m in self._metrics:
            at = np.ascontiguousarray(sorted(int(k) for k in (p["eval_at"] or (1, 2, 3, 4, 5))), np.int32)
            gain = None if p["label_gain"] is None else np.ascontiguousarray(p["label_gain"], np.float64)
            threads = int(p["num_threads"]) if int(p["num_threads"]) > 0 else 1
            o = _capi.MetricOpts(_capi.METRIC_IDS[m], int(p["multi_error_top_k"]), threads,
                                 len(at) if m == "ndcg" else 0, at.ctypes.data if m == "ndcg" else None,
                                 0 if gain is None else len(gain), None if gain is None else gain.ctypes.data)
            out = np.zeros(max(len(at), 1), np.float64)
            cnt = C.c_int32(0)
            check(lib().lgbmhip_eval(self._ctx, data_index, C.byref(o), ptr(out), C.byref(cnt)), self._ctx)
            if m == "ndcg":
                names = [f"ndcg@{k}" for k in at]
            elif m == "multi_error" and int(p["multi_error_top_k"]) != 1:
                names = [f"multi_error@{int(p['multi_error_top_k'])}"]
            else:
                names = [m]
            for nm, v in zip(names, out[: cnt.value]):
                res.append(EvalResult(data_name, nm, float(v), m in HIGHER_BETTER))
        return res

    def eval_train(self):
        return self._eval(0, self._train_name)

    def eval_valid(self):
        out = []
        for name, idx, _ in self._valid:
            out.extend(self._eval(idx, name))
        return out

    # --- prediction (basic.py:1641 -> py/lib.rs:557-579 -> predict.rs:231)
    def _window(self, start_iteration, num_iteration, pred_early_stop, pred_early_stop_freq,
                pred_early_stop_margin):
        # basic.py predict: num_iteration None means best_iteration when set, else every iteration
        if num_iteration is None:
            num_iteration = self.best_iteration if getattr(self, "best_iteration", 0) > 0 else -1
        freq = int(pred_early_stop_freq) if pred_early_stop else 0
        check(lib().lgbmhip_set_predict_window(self._ctx, int(start_iteration), int(num_iteration), freq,
                                               float(pred_early_stop_margin)), self._ctx)

    def predict(self, data, raw_score=False, start_iteration=0, num_iteration=None, pred_leaf=False,
                pred_early_stop=False, pred_early_stop_freq=10, pred_early_stop_margin=10.0):
        X = np.ascontiguousarray(data, np.float32)
        self._window(start_iteration, num_iteration, pred_early_stop, pred_early_stop_freq, pred_early_stop_margin)
        if pred_leaf:
            # PredictKind::LeafIndex: one column per tree of the window
            total = self.num_trees()
            out = np.zeros(max(X.shape[0] * total, 1), np.float64)
            nt = C.c_int32(0)
            check(lib().lgbmhip_predict_leaf(self._ctx, ptr(X), X.shape[0], X.shape[1], ptr(out), C.byref(nt)),
                  self._ctx)
            return out[: X.shape[0] * nt.value].reshape(X.shape[0], nt.value).astype(np.int32)
        out = np.zeros(X.shape[0] * self.K, np.float64)
        check(lib().lgbmhip_predict_raw(self._ctx, ptr(X), X.shape[0], X.shape[1], ptr(out)), self._ctx)
        raw = out.reshape(X.shape[0], self.K)
        if raw_score:
            return raw[:, 0] if self.K == 1 else raw
        obj = self.params["objective"]
        if obj == "binary":
            return 1.0 / (1.0 + np.exp(-float(self.params["sigmoid"]) * raw[:, 0]))
        if self.K > 1:
            e = np.exp(raw - raw.max(axis=1, keepdims=True))
            return e / e.sum(axis=1, keepdims=True)
        return raw[:, 0]

    def bin_rows_device(self, x_ptr, rows, ncol, out_ptr):
        """Bins rows resident in device memory (rows x ncol f32, integer device
        addresses) into device u8 [rows][F] with the training mappers."""
        check(lib().lgbmhip_bin_rows_device(self._ctx, x_ptr, rows, ncol, out_ptr), self._ctx)

    def predict_binned(self, bins, start_iteration=0, num_iteration=None):
        b = np.ascontiguousarray(bins, np.uint8)
        out = np.zeros(b.shape[0] * self.K, np.float64)
        self._window(start_iteration, num_iteration, False, 0, 0.0)
        check(lib().lgbmhip_predict_binned(self._ctx, ptr(b), b.shape[0], ptr(out)), self._ctx)
        return out.reshape(b.shape[0], self.K)

    def predict_device(self, x_ptr, rows, ncol, out_ptr, binned=False):
        """Raw scores of rows resident in device memory (integer device
        addresses, e.g. torch tensors' data_ptr()): x is rows x ncol f32, or
        rows x F u8 bins when `binned`; out is rows x K f64. With torch,
        import torch before this package first loads its library, so the
        process holds one HIP runtime."""
        self._window(0, None, False, 0, 0.0)
        if binned:
            check(lib().lgbmhip_predict_device(self._ctx, None, x_ptr, rows, 0, out_ptr), self._ctx)
        else:
            check(lib().lgbmhip_predict_device(self._ctx, x_ptr, None, rows, ncol, out_ptr), self._ctx)

    def profile(self, enable=-1):
        """Kernel-kind timing totals (see lgbmhip_profile); enable=1 starts, 0 stops, -1 reads."""
        out = np.zeros(12, np.float64)
        check(lib().lgbmhip_profile(self._ctx, int(enable), ptr(out)), self._ctx)
        keys = ("hist_ms", "hist_launches", "hist_rows", "hist_bytes", "split_ms", "part_ms", "iter_ms", "iters",
                "part_rows", "splits_computed", "splits_used", "slab_ms")
        return dict(zip(keys, out.tolist()))

    def lds_atomic_peak(self):
        """Lane atomics per second of conflict-free ds_add_u32 on this device (lgbmhip_lds_atomic_peak)."""
        out = np.zeros(1, np.float64)
        check(lib().lgbmhip_lds_atomic_peak(self._ctx, ptr(out)), self._ctx)
        return float(out[0])

    def set_tuning(self, part_grid=0, hist_chunks=0, hist_min_rows=0, hist_lds_budget=0, rounds_main=0,
                   rounds_cont=0, late_round=-1):
        """Launch shapes of the tree-growth kernels (before the first update;
        0 keeps a value, and -1 for late_round, where 0 means never). Trees
        do not depend on them."""
        check(lib().lgbmhip_set_tuning(self._ctx, int(part_grid), int(hist_chunks), int(hist_min_rows),
                                       int(hist_lds_budget), int(rounds_main), int(rounds_cont), int(late_round)),
              self._ctx)

    def bin_layout(self):
        """The training rows in HBM: bytes per packed row, its padded stride,
        and how many features are stored as 4-bit fields."""
        out = np.zeros(3, np.int32)
        check(lib().lgbmhip_bin_layout(self._ctx, ptr(out)), self._ctx)
        return dict(zip(("row_bytes", "row_stride", "u4_features"), out.tolist()))

    def get_tuning(self):
        out = np.zeros(10, np.int32)
        check(lib().lgbmhip_get_tuning(self._ctx, ptr(out)), self._ctx)
        keys = ("part_grid", "hist_chunks", "hist_min_rows", "hist_lds", "hist_features_per_block",
                "hist_groups", "hist_threads_per_row", "rounds_main", "rounds_cont", "late_round")
        return dict(zip(keys, out.tolist()))

    def last_iteration_ms(self):
        out = np.zeros(4, np.float64)
        check(lib().lgbmhip_last_timings(self._ctx, ptr(out)), self._ctx)
        return float(out[0])

    def free(self):
        if getattr(self, "_ctx", None) and _capi._lib is not None:
            _capi._lib.lgbmhip_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        self.free()


def train(params, train_set, num_boost_round=100, valid_sets=None, valid_names=None, callbacks=None):
    """engine.py:50-200: the boosting loop, evaluation of the validation sets
    after each iteration, and early stopping (callback.py:161-305 semantics:
    a validation metric without improvement for early_stopping_round
    iterations stops training; best_iteration is 1-based)."""
    p = _params(params)
    n = int(p.get("num_iterations", num_boost_round)) if "num_iterations" in (params or {}) else num_boost_round
    if n <= 0:
        raise ValueError(f"Number of boosting rounds must be greater than 0. Got {n}.")
    b = Booster(params, train_set)
    if isinstance(valid_sets, Dataset):
        valid_sets = [valid_sets]
    if isinstance(valid_names, str):
        valid_names = [valid_names]
    eval_train = False
    for i, v in enumerate(valid_sets or []):
        name = valid_names[i] if valid_names is not None and len(valid_names) > i else f"valid_{i}"
        if v is train_set:
            eval_train = True
            b.set_train_data_name(name)
            continue
        b.add_valid(v, name)
    rounds = int(p["early_stopping_round"] or 0)
    first_only = _b(p["first_metric_only"])
    delta = float(p["early_stopping_min_delta"])
    best, best_iter, best_list = [], [], []
    results = []
    for i in range(n):
        if b.update():
            break
        results = []
        if valid_sets is not None:
            if eval_train:
                results.extend(b.eval_train())
            results.extend(b.eval_valid())
        for cb in callbacks or []:
            cb(CallbackEnv(b, p, i, 0, n, results))
        if rounds <= 0 or not results:
            continue
        if not best:
            best = [-np.inf if r.maximize else np.inf for r in results]
            best_iter = [0] * len(results)
            best_list = [results] * len(results)
        stop = None
        for j, r in enumerate(results):
            better = r.metric_value > best[j] + delta if r.maximize else r.metric_value < best[j] - delta
            if better or i == 0:
                best[j], best_iter[j], best_list[j] = r.metric_value, i, results
            if first_only and r.metric_name != results[0].metric_name:
                continue
            if r.dataset_name == b._train_name and eval_train:
                continue
            if i - best_iter[j] >= rounds or i == n - 1:
                stop = j
                break
        if stop is not None:
            b.best_iteration = best_iter[stop] + 1
            results = best_list[stop]
            break
    b.best_score = {}
    for r in results:
        b.best_score.setdefault(r.dataset_name, {})[r.metric_name] = r.metric_value
    return b
