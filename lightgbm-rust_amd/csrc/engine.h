// Shared declarations of the MI355X GBDT engine: host binning types, the
// device-side records of the tree-growth state machine, and kernel launchers.
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/lgbm_hip.h"

namespace lgbmhip {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// One column's bin mapper (binning.rs:43-64).
struct Mapper {
  int num_bin = 1;
  int missing_type = 0;
  std::vector<double> bounds{INFINITY};
  bool is_trivial = true;
  double sparse_rate = 1.0;
  int bin_type = 0;
  uint32_t default_bin = 0, most_freq_bin = 0;
  std::vector<int> categories;
  std::unordered_map<int, int> cat_to_bin;
  uint32_t to_bin(double v) const;
};

// Result of host binning: mappers of every column, the used columns and the
// n x F row-major u8 bin matrix.
struct Binned {
  int64_t n = 0;
  int ncol = 0;
  std::vector<Mapper> mappers;
  std::vector<int> used;
  std::vector<uint8_t> bins;
};

Binned* bin_dense(const float* X, int64_t n, int ncol, const int32_t* cats, int n_cat, const lgbmhip_bin_params& p);
void apply_bins(const Binned& b, const float* X, int64_t n, int ncol, uint8_t* out);

// ---------------------------------------------------------------- device records
constexpr int kMaxBins = 256;        // u8 bins
constexpr int kCatWords = kMaxBins / 32;

// Per used feature, as the split search and partition read it (split.rs:18-29).
struct FeatDev {
  int num_bin, missing_type, default_bin, mfb, bin_type, offset;
};

// Split parameters in device form (split.rs:33-46).
struct SplitParamsDev {
  double lambda_l1, lambda_l2, max_delta_step, path_smooth;
  double min_sum_hessian, min_gain_to_split, cat_l2, cat_smooth;
  int min_data_in_leaf, max_cat_to_onehot, max_cat_threshold, min_data_per_group;
  int num_leaves, max_depth;
};

// A split candidate or a chosen split (split.rs:60-84 SplitInfo).
struct SplitRec {
  double gain, left_output, right_output, lsg, lsh, rsg, rsh;
  long long lgh, rgh;
  int feature;      // inner feature, -1 none
  int threshold;
  int default_left;
  int left_count, right_count;
  int num_cat;      // categorical: number of bins sent left
  int leaf;         // split log: the leaf this split was applied to
  unsigned cat_bits[kCatWords];  // categorical: bins sent left, as a bitset over bins
};

// One leaf of the tree being grown (partition.rs segment + leaf_splits.hpp sums).
struct LeafDev {
  long long begin, count;
  int buf, depth;
  double sum_g, sum_h, output;
  long long int_sum;
};

// The tree-growth state machine, one record per context.
struct StepDev {
  int num_leaves;   // leaves of the tree so far
  int done;         // no further split this tree
  int cur_leaf;     // leaf split in this step
  int new_leaf;
  int smaller, larger;  // leaves whose histograms / searches run next (larger = -1 at the root)
  int skip;         // before_find_best_split said no
  int is_root;
  long long left_total;
  int hist_shift_g, hist_shift_h;  // fixed-point exponents of the current histogram build
  int num_splits;
  int counter;      // last-block-done counter
  int counter2;
  double root_output;  // the root's output before any split (learner/mod.rs:647)
  long long seg_begin, seg_count;  // the segment of the leaf being split
  int seg_buf, seg_depth;
  unsigned long long lcur, rcur;   // rows placed left / right by the partition pass
};

// Per-iteration scalars written by device reductions.
struct IterDev {
  unsigned max_abs_g_bits, max_abs_h_bits;  // float bits of max |g|, |h| (non-negative floats order as ints)
  long long root_g_fix, root_h_fix;         // fixed-point root sums
  long long root_int_sum;                   // packed int8 root sum (quantized form)
  long long root_gq_sum, root_hq_sum;
  double grad_scale, hess_scale;            // quantizer scales
  int start;                                // stochastic-rounding start offset
  int pad;
};

}  // namespace lgbmhip
