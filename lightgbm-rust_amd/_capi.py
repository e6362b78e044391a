"""ctypes binding of liblgbm_hip.so (the C ABI declared in include/lgbm_hip.h).

The library is required: there is no CPU fallback. Importing this module on a
tree whose library was not built raises, and every entry point that fails
raises LightGBMError with the library's message.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("LGBMHIP_LIB") or os.path.join(_HERE, "_lib", "liblgbm_hip.so")

OK = 0
STATUS_NAMES = {1: "InvalidParameter", 2: "InvalidData", 3: "ModelFormat", 4: "Unsupported", 5: "Internal"}


class LightGBMError(Exception):
    """Mirrors the reference binding's LightGBMError (crates/lgbm-py/src/lib.rs:89-100)."""

    def __init__(self, msg, status=None):
        super().__init__(msg)
        self.status = status


class BinParams(C.Structure):
    _fields_ = [("max_bin", C.c_int32), ("min_data_in_bin", C.c_int32), ("min_data_in_leaf", C.c_int32),
                ("bin_construct_sample_cnt", C.c_int32), ("data_random_seed", C.c_int32),
                ("feature_pre_filter", C.c_int32), ("use_missing", C.c_int32), ("zero_as_missing", C.c_int32),
                ("num_threads", C.c_int32)]


class FeatureMeta(C.Structure):
    _fields_ = [("real_index", C.c_int32), ("num_bin", C.c_int32), ("missing_type", C.c_int32),
                ("default_bin", C.c_int32), ("most_freq_bin", C.c_int32), ("bin_type", C.c_int32)]


class SplitParams(C.Structure):
    _fields_ = [("lambda_l1", C.c_double), ("lambda_l2", C.c_double), ("max_delta_step", C.c_double),
                ("path_smooth", C.c_double), ("min_sum_hessian_in_leaf", C.c_double),
                ("min_gain_to_split", C.c_double), ("cat_l2", C.c_double), ("cat_smooth", C.c_double),
                ("min_data_in_leaf", C.c_int32), ("max_cat_to_onehot", C.c_int32), ("max_cat_threshold", C.c_int32),
                ("min_data_per_group", C.c_int32), ("num_leaves", C.c_int32), ("max_depth", C.c_int32)]


class TrainOpts(C.Structure):
    _fields_ = [("objective", C.c_int32), ("num_class", C.c_int32), ("sigmoid", C.c_double),
                ("learning_rate", C.c_double), ("boost_from_average", C.c_int32),
                ("use_quantized_grad", C.c_int32), ("num_grad_quant_bins", C.c_int32),
                ("stochastic_rounding", C.c_int32), ("quant_train_renew_leaf", C.c_int32), ("seed", C.c_int32),
                ("quant_table_threads", C.c_int32), ("lambdarank_truncation_level", C.c_int32),
                ("lambdarank_norm", C.c_int32), ("use_graph", C.c_int32), ("bagging_fraction", C.c_double),
                ("bagging_freq", C.c_int32), ("bagging_seed", C.c_int32), ("pos_bagging_fraction", C.c_double),
                ("neg_bagging_fraction", C.c_double), ("data_sample_strategy", C.c_int32),
                ("sample_threads", C.c_int32), ("top_rate", C.c_double), ("other_rate", C.c_double),
                ("bagging_by_query", C.c_int32), ("batch_trees", C.c_int32)]


class MetricOpts(C.Structure):
    _fields_ = [("metric", C.c_int32), ("multi_error_top_k", C.c_int32), ("num_threads", C.c_int32),
                ("num_eval_at", C.c_int32), ("eval_at", C.c_void_p), ("num_label_gain", C.c_int32),
                ("label_gain", C.c_void_p)]


METRIC_IDS = {"binary_logloss": 0, "binary_error": 1, "auc": 2, "multi_logloss": 3, "multi_error": 4, "ndcg": 5}

P = C.c_void_p
I32 = C.c_int32
I64 = C.c_int64
D = C.c_double
PP = C.POINTER(C.c_void_p)

# name -> (restype, argtypes); every declaration of include/lgbm_hip.h
SIGNATURES = {
    "lgbmhip_last_error": (C.c_char_p, [P]),
    "lgbmhip_global_last_error": (C.c_char_p, []),
    "lgbmhip_bin_dense": (I32, [P, I64, I32, P, I32, C.POINTER(BinParams), PP]),
    "lgbmhip_find_mappers": (I32, [P, I64, I32, P, I32, C.POINTER(BinParams), PP]),
    "lgbmhip_bins_free": (None, [P]),
    "lgbmhip_bins_num_features": (I32, [P]),
    "lgbmhip_bins_num_groups": (I32, [P]),
    "lgbmhip_bins_metas": (I32, [P, P, P]),
    "lgbmhip_bins_num_upper": (I64, [P]),
    "lgbmhip_bins_matrix": (I32, [P, P]),
    "lgbmhip_bins_apply": (I32, [P, P, I64, I32, P]),
    "lgbmhip_create": (I32, [I32, PP]),
    "lgbmhip_destroy": (None, [P]),
    "lgbmhip_upload_bins": (I32, [P, P, I32, P, P, I64, I32]),
    "lgbmhip_upload_raw": (I32, [P, P, P, I64, I32, I32]),
    "lgbmhip_bin_rows_device": (I32, [P, P, I64, I32, P]),
    "lgbmhip_set_metadata": (I32, [P, P, P, P, I32]),
    "lgbmhip_set_split_params": (I32, [P, C.POINTER(SplitParams)]),
    "lgbmhip_set_train_opts": (I32, [P, C.POINTER(TrainOpts)]),
    "lgbmhip_train_iteration": (I32, [P, C.POINTER(I32)]),
    "lgbmhip_train_tree": (I32, [P, P, P]),
    "lgbmhip_sync": (I32, [P]),
    "lgbmhip_num_trees": (I32, [P]),
    "lgbmhip_tree_num_leaves": (I32, [P, I32]),
    "lgbmhip_tree_arrays": (I32, [P, I32, P, P, P, P, P, P, P, P, P, P, P]),
    "lgbmhip_leaf_of_row": (I32, [P, P]),
    "lgbmhip_add_leaf_outputs": (I32, [P, P, I32, P]),
    "lgbmhip_copy_scores": (I32, [P, P]),
    "lgbmhip_copy_gradients": (I32, [P, P, P]),
    "lgbmhip_quant_state": (I32, [P, P, P, P]),
    "lgbmhip_gradients": (I32, [P, P, P, P]),
    "lgbmhip_discretize": (I32, [P, P, P, I32, P, P, P]),
    "lgbmhip_build_histogram": (I32, [P, P, I64, P, P, P, P]),
    "lgbmhip_build_histogram_int": (I32, [P, P, I64, P, P, P, P]),
    "lgbmhip_find_best_split": (I32, [P, I32, P, D, D, I32, D, P, P, P]),
    "lgbmhip_find_best_split_int": (I32, [P, I32, P, I64, D, D, I32, D, P, P, P, P]),
    "lgbmhip_partition": (I32, [P, P, I64, I32, C.c_uint32, I32, P, P]),
    "lgbmhip_partition_device_noalloc": (I32, [P, P, I64, I32, C.c_uint32, I32, P, C.POINTER(I64)]),
    "lgbmhip_predict_raw": (I32, [P, P, I64, I32, P]),
    "lgbmhip_predict_binned": (I32, [P, P, I64, P]),
    "lgbmhip_predict_device": (I32, [P, P, P, I64, I32, P]),
    "lgbmhip_last_timings": (I32, [P, P]),
    "lgbmhip_add_valid": (I32, [P, P, I64, P, P, P, I32, P, C.POINTER(I32)]),
    "lgbmhip_copy_valid_scores": (I32, [P, I32, P]),
    "lgbmhip_eval": (I32, [P, I32, C.POINTER(MetricOpts), P, C.POINTER(I32)]),
    "lgbmhip_set_predict_window": (I32, [P, I32, I32, I32, D]),
    "lgbmhip_predict_leaf": (I32, [P, P, I64, I32, P, C.POINTER(I32)]),
    "lgbmhip_rollback_one_iter": (I32, [P]),
    "lgbmhip_save_model": (I32, [P, I32, C.c_char_p, C.c_char_p, C.c_char_p, P, I64, C.POINTER(I64)]),
    "lgbmhip_load_model": (I32, [P, C.c_char_p]),
    "lgbmhip_check_model": (I32, [C.c_char_p, P, P, P, I64, C.POINTER(I64)]),
    "lgbmhip_bins_feature_infos": (I32, [P, P, I64, C.POINTER(I64)]),
    "lgbmhip_profile": (I32, [P, I32, P]),
    "lgbmhip_lds_atomic_peak": (I32, [P, P]),
    "lgbmhip_set_tuning": (I32, [P, I32, I32, I32, I32, I32, I32, I32]),
    "lgbmhip_autotune": (I32, [P, I32, P]),
    "lgbmhip_get_tuning": (I32, [P, P]),
    "lgbmhip_set_bin_packing": (I32, [P, I32]),
    "lgbmhip_set_feature_groups": (I32, [P, I32]),
    "lgbmhip_bin_layout": (I32, [P, P]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built: run `python lightgbm-rust_amd/build.py` "
                              "(the HIP engine has no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def check(status, ctx=None):
    if status != OK:
        msg = lib().lgbmhip_last_error(ctx) if ctx else lib().lgbmhip_global_last_error()
        msg = msg.decode() if msg else ""
        raise LightGBMError(f"{STATUS_NAMES.get(status, status)}: {msg}", status)
