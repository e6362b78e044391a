"""Tree comparison shared by the GPU parity tests: the engine's trees against
the oracle's, exactly, except at a reported near-tie."""
import numpy as np


def first_difference(gt, ot):
    """Index of the first internal node whose split differs, or None."""
    n = min(gt.num_leaves, ot.num_leaves) - 1
    for k in range(n):
        if (gt.split_feature[k] != ot.split_feature[k] or gt.threshold_in_bin[k] != ot.threshold_in_bin[k]
                or gt.left_child[k] != ot.left_child[k] or gt.right_child[k] != ot.right_child[k]
                or gt.decision_type[k] != ot.decision_type[k]):
            return k
    return None if gt.num_leaves == ot.num_leaves else n


K_ZERO = float(np.float32(1e-35))


def boosted_values(v, lr, bias):
    """Tree::shrink then Tree::add_bias (tree.rs:303-334), each rounding
    |x| <= kZeroThreshold to 0: the leaf values a boosting iteration keeps."""
    v = v * lr
    v = np.where(np.abs(v) <= K_ZERO, 0.0, v)
    if bias != 0.0:
        v = v + bias
        v = np.where(np.abs(v) <= K_ZERO, 0.0, v)
    return v


def compare_or_near_tie(gt, ot, lr=1.0, bias=0.0, value_rtol=1e-9):
    """True when the trees are equal; False after an accepted near-tie. The
    oracle's tree is the learner's (unshrunk); the GPU's is the booster's."""
    k = first_difference(gt, ot)
    if k is None:
        np.testing.assert_array_equal(gt.threshold, ot.threshold)
        np.testing.assert_array_equal(gt.leaf_count, ot.leaf_count)
        np.testing.assert_allclose(gt.leaf_value, boosted_values(ot.leaf_value, lr, bias), rtol=value_rtol,
                                   atol=1e-12)
        np.testing.assert_allclose(gt.split_gain, ot.split_gain, rtol=1e-6)
        return True
    # nodes are numbered in split order: every split before k is the same, and
    # split k's two choices must have the same gain at f32 resolution
    assert k < min(gt.num_leaves, ot.num_leaves) - 1, "tree sizes differ with equal splits"
    print(f"[trees] near-tie at split {k}: gpu feature {gt.split_feature[k]} bin {gt.threshold_in_bin[k]} "
          f"gain {gt.split_gain[k]!r}, oracle feature {ot.split_feature[k]} bin {ot.threshold_in_bin[k]} "
          f"gain {ot.split_gain[k]!r}", flush=True)
    assert gt.split_gain[k] == ot.split_gain[k], (k, gt.split_gain[k], ot.split_gain[k])
    return False
