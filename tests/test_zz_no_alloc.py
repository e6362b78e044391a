"""lgbmhip_partition_device_noalloc: the stable partition of device rows with
its work buffer in a static per-device array (no device allocation), checked
against lgbmhip_partition (the allocating entry point over host rows, itself
checked against numpy's stable partition) on consecutive calls, the second
smaller than the first, so stale contents of the shared buffer would show."""
import ctypes as C

import numpy as np
import pytest

from tests._pkg import DeviceArray, capi, lgb, small_binary

pytestmark = pytest.mark.gpu

MAX_ROWS = (1 << 16) - 2
MAX_ROWS *= 1024


def _host_partition(b, rows, f, thr, dl):
    out = np.zeros_like(rows)
    nl = np.zeros(1, np.int64)
    capi.check(capi.lib().lgbmhip_partition(b._ctx, capi.ptr(rows), len(rows), f, thr, dl, capi.ptr(out),
                                            capi.ptr(nl)), b._ctx)
    return out, int(nl[0])


def _noalloc_partition(b, rows, f, thr, dl):
    d_rows = DeviceArray(rows)
    d_out = DeviceArray(nbytes=max(rows.nbytes, 4))
    nl = C.c_int64(-1)
    capi.check(capi.lib().lgbmhip_partition_device_noalloc(b._ctx, C.c_void_p(d_rows.ptr), len(rows), f, thr, dl,
                                                           C.c_void_p(d_out.ptr), C.byref(nl)), b._ctx)
    out = np.zeros_like(rows)
    if len(rows):
        d_out.to_host(out)
    return out, nl.value


@pytest.fixture(scope="module")
def booster():
    X, y = small_binary(300000, 5, nan_col=2, seed=11)
    params = dict(objective="binary", num_leaves=31, max_bin=255)
    ds = lgb.Dataset(X, label=y, params=params)
    return ds, lgb.Booster(params, ds)


def test_noalloc_partition_equals_the_allocating_entry_point(booster):
    ds, b = booster
    rng = np.random.default_rng(12)
    n = ds.num_data
    # two calls in a row per case, the second smaller (the buffer keeps the first's tile counts)
    for f in range(ds.num_features):
        meta = ds.feature_meta(f)
        thr = int(rng.integers(0, meta["num_bin"]))
        for sizes in ((n, 4097), (200000, 1), (70001, 0)):
            for m in sizes:
                rows = rng.choice(n, m, replace=False).astype(np.uint32)   # any order; stable in it
                for dl in (0, 1):
                    want, wl = _host_partition(b, rows, f, thr, dl)
                    got, gl = _noalloc_partition(b, rows, f, thr, dl)
                    assert gl == wl, (f, m, dl)
                    np.testing.assert_array_equal(got, want)
                    bins = ds.bins[rows, f].astype(int)
                    miss = (meta["missing_type"] == 2) & (bins == meta["num_bin"] - 1)
                    left = np.where(miss, bool(dl), bins <= thr)
                    np.testing.assert_array_equal(want, np.concatenate([rows[left], rows[~left]]))


def test_noalloc_partition_rejects_bad_input_without_faulting(booster):
    ds, b = booster
    lib = capi.lib()
    nl = C.c_int64(-1)
    # a row index past the data: the kernels read nothing for it and the call reports InvalidData
    rows = np.arange(5000, dtype=np.uint32)
    rows[4321] = ds.num_data + 7
    d_rows, d_out = DeviceArray(rows), DeviceArray(nbytes=rows.nbytes)
    st = lib.lgbmhip_partition_device_noalloc(b._ctx, C.c_void_p(d_rows.ptr), len(rows), 0, 3, 0,
                                              C.c_void_p(d_out.ptr), C.byref(nl))
    assert st == 2, capi.lib().lgbmhip_last_error(b._ctx)
    # more rows than the buffer serves: refused before any launch
    st = lib.lgbmhip_partition_device_noalloc(b._ctx, C.c_void_p(d_rows.ptr), MAX_ROWS + 1, 0, 3, 0,
                                              C.c_void_p(d_out.ptr), C.byref(nl))
    assert st == 1 and b"at most" in lib.lgbmhip_last_error(b._ctx)
    # and a good call after the errors is still right
    good = np.arange(0, ds.num_data, 3, dtype=np.uint32)
    want, wl = _host_partition(b, good, 1, 10, 1)
    got, gl = _noalloc_partition(b, good, 1, 10, 1)
    assert gl == wl
    np.testing.assert_array_equal(got, want)
