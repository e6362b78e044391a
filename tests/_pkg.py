"""Imports the engine package (its directory name carries a hyphen)."""
import importlib
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

lgb = importlib.import_module("lightgbm-rust_amd")
datagen = importlib.import_module("lightgbm-rust_amd.datagen")
capi = importlib.import_module("lightgbm-rust_amd._capi")


def small_binary(n=5000, F=8, seed=0, nan_col=None, zero_col=None):
    import numpy as np
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, F)).astype(np.float32)
    if nan_col is not None:
        X[::7, nan_col] = np.nan
    if zero_col is not None:
        X[::3, zero_col] = 0.0
    z = X[:, 0] - 0.5 * np.nan_to_num(X[:, 1]) ** 2 + 0.3 * rng.normal(size=n)
    y = (z > 0).astype(np.float32)
    return X, y
