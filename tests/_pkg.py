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


class DeviceArray:
    """A device copy of a numpy array through the HIP runtime the engine
    library already loaded (hipMalloc / hipMemcpy via ctypes), for entry
    points that take device pointers."""

    def __init__(self, host=None, nbytes=None):
        import ctypes as C
        capi.lib()   # the engine's HIP runtime is the one in the process
        self._hip = C.CDLL("libamdhip64.so.7")
        self.nbytes = int(host.nbytes if host is not None else nbytes)
        p = C.c_void_p()
        assert self._hip.hipMalloc(C.byref(p), C.c_size_t(max(self.nbytes, 1))) == 0
        self.ptr = p.value
        if host is not None:
            assert self._hip.hipMemcpy(C.c_void_p(self.ptr), host.ctypes.data_as(C.c_void_p),
                                       C.c_size_t(self.nbytes), 1) == 0   # hipMemcpyHostToDevice

    def to_host(self, out):
        import ctypes as C
        assert out.nbytes <= self.nbytes
        assert self._hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), C.c_size_t(out.nbytes),
                                   2) == 0   # hipMemcpyDeviceToHost
        return out

    def __del__(self):
        import ctypes as C
        if getattr(self, "ptr", None):
            self._hip.hipFree(C.c_void_p(self.ptr))
            self.ptr = None
