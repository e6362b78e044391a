"""Model text (io.rs:13-343, tree.rs:718-770): the engine's trees written as
the reference's text format and read back for raw prediction on the GPU.
Numbers are written with C's %.17g, so the round trip is exact: predictions
of a loaded model equal the trained booster's bit for bit. The parameters
section is not written (it is the reference's whole config dump)."""
import numpy as np
import pytest

from tests._pkg import datagen, lgb

pytestmark = pytest.mark.gpu


def _blocks(text):
    head, rest = text.split("\n\n", 1)
    trees, tail = rest.split("end of trees\n", 1)
    return head, trees, tail


@pytest.mark.parametrize("objective", ["binary", "multiclass"])
def test_save_load_round_trip(objective):
    X, y = datagen.binary_100k(30000)
    X = X.copy()
    X[:, 3] = np.floor(np.abs(X[:, 3]) * 3)   # a categorical column
    X[::9, 2] = np.nan
    params = dict(objective=objective, num_leaves=31, max_bin=63, categorical_feature=[3])
    if objective == "multiclass":
        y = np.minimum(np.floor(np.abs(X[:, 0]) * 2), 2).astype(np.float32)
        params["num_class"] = 3
    ds = lgb.Dataset(X, label=y, params=params)
    b = lgb.Booster(params, ds)
    for _ in range(5):
        b.update()
    text = b.model_to_string()
    head, trees, tail = _blocks(text)
    kv = dict(l.split("=", 1) for l in head.splitlines() if "=" in l)
    assert head.startswith("tree\nversion=v4\n")
    assert int(kv["num_tree_per_iteration"]) == b.K and int(kv["max_feature_idx"]) == X.shape[1] - 1
    assert kv["objective"] == ("binary sigmoid:1" if objective == "binary" else "multiclass num_class:3")
    assert len(kv["feature_infos"].split(" ")) == X.shape[1]
    # tree_sizes are the byte lengths of the "Tree=i" blocks
    sizes = [int(v) for v in kv["tree_sizes"].split(" ")]
    assert len(sizes) == b.num_trees()
    off = 0
    for i, sz in enumerate(sizes):
        blk = trees[off:off + sz]
        assert blk.startswith(f"Tree={i}\nnum_leaves=") and blk.endswith("shrinkage=1\n\n\n") or blk.endswith("\n\n\n")
        lv = [float(v) for v in next(l for l in blk.splitlines() if l.startswith("leaf_value=")).split("=")[1].split()]
        assert lv == b.tree(i).leaf_value.tolist()   # %.17g round-trips every double
        off += sz
    assert off == len(trees)
    assert tail.startswith("\nfeature_importances:\n")
    loaded = lgb.Booster(model_str=text)
    Xt = X[:5000]
    np.testing.assert_array_equal(loaded.predict(Xt, raw_score=True), b.predict(Xt, raw_score=True))


def test_hand_written_model_known_answers():
    # two trees: a numerical split with default-left NaN handling, and a
    # categorical split on categories {1, 3}
    text = "\n".join([
        "tree", "version=v4", "num_class=1", "num_tree_per_iteration=1", "label_index=0", "max_feature_idx=2",
        "objective=binary sigmoid:1", "feature_names=a b c", "feature_infos=none none none", "tree_sizes=0 0", "",
        "Tree=0", "num_leaves=2", "num_cat=0", "split_feature=1", "split_gain=1", "threshold=0.5",
        "decision_type=10", "left_child=-1", "right_child=-2", "leaf_value=-1 2", "leaf_weight=1 1",
        "leaf_count=1 1", "internal_value=0", "internal_weight=0", "internal_count=2", "is_linear=0",
        "shrinkage=1", "", "",
        "Tree=1", "num_leaves=2", "num_cat=1", "split_feature=2", "split_gain=1", "threshold=0",
        "decision_type=1", "left_child=-1", "right_child=-2", "leaf_value=0.25 0.5", "leaf_weight=1 1",
        "leaf_count=1 1", "internal_value=0", "internal_weight=0", "internal_count=2", "cat_boundaries=0 1",
        "cat_threshold=10", "is_linear=0", "shrinkage=1", "", "", "end of trees", ""])
    b = lgb.Booster(model_str=text)
    X = np.array([[0, 0.5, 1], [0, 0.6, 3], [0, np.nan, 2], [0, 7.0, np.nan], [0, -1.0, -3]], np.float32)
    # tree 0: decision_type 10 = default_left (2) | NaN missing (8): x <= 0.5 or NaN go left
    # tree 1: categories 1 and 3 (bitset 0b1010) go left; NaN and negatives go right
    want = np.array([-1 + 0.25, 2 + 0.25, -1 + 0.5, 2 + 0.5, -1 + 0.5])
    np.testing.assert_array_equal(b.predict(X, raw_score=True), want)


def test_two_categorical_bitsets_known_answers():
    # one tree with two categorical nodes whose bitsets differ: node 0 sends
    # categories {1, 3} of column 0 left (bitset 0: 0b1010), node 1 sends
    # category {2} of column 1 left (bitset 1: 0b100). A node's bitset is the
    # one its threshold names (tree.rs:400-403).
    text = "\n".join([
        "tree", "version=v4", "num_class=1", "num_tree_per_iteration=1", "label_index=0", "max_feature_idx=1",
        "objective=binary sigmoid:1", "feature_names=a b", "feature_infos=none none", "",
        "Tree=0", "num_leaves=3", "num_cat=2", "split_feature=0 1", "split_gain=1 1", "threshold=0 1",
        "decision_type=1 1", "left_child=-1 -2", "right_child=1 -3", "leaf_value=10 20 30",
        "leaf_weight=1 1 1", "leaf_count=1 1 1", "internal_value=0 0", "internal_weight=0 0",
        "internal_count=3 2", "cat_boundaries=0 1 2", "cat_threshold=10 4", "is_linear=0", "shrinkage=1", "", "",
        "end of trees", ""])
    b = lgb.Booster(model_str=text)
    X = np.array([[1, 0], [3, 2], [0, 2], [2, 1], [5, 2], [0, 3], [np.nan, 2]], np.float32)
    want = np.array([10, 10, 20, 30, 20, 30, 20], np.float64)
    np.testing.assert_array_equal(b.predict(X, raw_score=True), want)


def test_save_of_load_of_save_is_stable():
    # tests/end_to_end.rs:59-61: save(load(save(trained))) equals save(trained)
    X, y = datagen.binary_100k(20000)
    X = X.copy()
    X[:, 5] = np.floor(np.abs(X[:, 5]) * 4)
    params = dict(objective="binary", num_leaves=15, max_bin=63, categorical_feature=[5])
    b = lgb.Booster(params, lgb.Dataset(X, label=y, params=params))
    for _ in range(4):
        b.update()
    block = "[boosting: gbdt]\n[objective: binary]\n[device_type: cpu]\n[num_leaves: 15]\n"
    text = b.model_to_string(parameters=block)
    assert text.endswith("\nparameters:\n" + block + "\nend of parameters\n")
    loaded = lgb.Booster(model_str=text)
    assert loaded.model_to_string() == text
    np.testing.assert_array_equal(loaded.predict(X[:3000], raw_score=True), b.predict(X[:3000], raw_score=True))
