"""The autotuner (lgbmhip_autotune): launch shapes timed on the workload's own
first trees. Trees do not depend on launch shapes, so a tuned booster trains
the untuned one's model bit for bit; the choice is cached per process for
the workload shape."""
import numpy as np
import pytest

from tests._pkg import datagen, lgb

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("quant", [False, True])
def test_autotuned_booster_trains_the_same_model(quant):
    X, y = datagen.binary_100k(200000)
    params = dict(objective="binary", num_leaves=63, max_bin=255, min_data_in_leaf=20, learning_rate=0.1)
    if quant:
        params.update(use_quantized_grad=True, num_grad_quant_bins=16, quant_train_renew_leaf=True)
    ds = lgb.Dataset(X, label=y, params=params)
    tuned = lgb.Booster(dict(params, autotune=True, autotune_trees=1), ds)
    t = tuned.tuning
    print(f"autotune quant={quant}: {t}", flush=True)
    assert t["candidates_timed"] >= 4
    got = tuned.get_tuning()
    assert got["part_grid"] == t["part_grid"] and got["hist_chunks"] == t["hist_chunks"]
    plain = lgb.Booster(params, ds)
    for _ in range(5):
        tuned.update()
        plain.update()
    np.testing.assert_array_equal(tuned.train_scores(), plain.train_scores())
    for i in range(5):
        a, b = tuned.tree(i), plain.tree(i)
        np.testing.assert_array_equal(a.split_feature, b.split_feature)
        np.testing.assert_array_equal(a.threshold_in_bin, b.threshold_in_bin)
        np.testing.assert_array_equal(a.leaf_value, b.leaf_value)
    # the same shape again: the cached choice, nothing timed
    again = lgb.Booster(dict(params, autotune=True), ds)
    assert again.tuning["candidates_timed"] == 0
    assert {k: again.tuning[k] for k in ("part_grid", "hist_chunks", "hist_min_rows", "hist_lds")} == \
        {k: t[k] for k in ("part_grid", "hist_chunks", "hist_min_rows", "hist_lds")}
    # only before the first iteration
    lib = lgb._capi.lib()
    assert lib.lgbmhip_autotune(tuned._ctx, 1, None) == 1
