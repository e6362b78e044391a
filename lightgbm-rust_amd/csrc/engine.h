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
  double min_val = 0.0, max_val = 0.0;   // smallest / largest distinct sample value (binning.rs:453-454)
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
  int num_groups = 0;   // the reference's feature groups (exclusive feature bundling) over `used`
};

Binned* bin_dense(const float* X, int64_t n, int ncol, const int32_t* cats, int n_cat, const lgbmhip_bin_params& p);
// The mappers only (no bin matrix): rows are binned on the device.
Binned* find_mappers(const float* X, int64_t n, int ncol, const int32_t* cats, int n_cat, const lgbmhip_bin_params& p);
// Rust's `f64 as i32`: saturating, NaN to 0 (binning.rs:641).
inline int sat_i32(double v) {
  if (!(v == v)) return 0;
  if (v >= 2147483647.0) return 2147483647;
  if (v <= -2147483648.0) return -2147483647 - 1;
  return (int)v;
}
void apply_bins(const Binned& b, const float* X, int64_t n, int ncol, uint8_t* out);

// ---------------------------------------------------------------- device records
constexpr int kMaxBins = 256;        // u8 bins
constexpr int kCatWords = kMaxBins / 32;

// Per used feature, what value_to_bin reads on the device (binning.rs:637-667):
// its raw column, and at `off` its bin upper bounds (numerical) or its
// categories sorted ascending with their bins (categorical, ncat of them).
struct BinFeat {
  int col, num_bin, missing_type, bin_type, off, ncat;
};

// Per used feature, as the split search and partition read it (split.rs:18-29),
// and where its bins live: byte `pos` of a packed row (nib -1: the whole
// byte; 0 / 1: the low / high 4 bits of a feature with at most 16 bins), and
// its histogram's first entry `hoff` in the histogram unit of group `hg`.
struct FeatDev {
  int num_bin, missing_type, default_bin, mfb, bin_type, offset;
  int pos, nib, hg, hoff;
};

// Byte p of a packed row as the histogram kernels read it: the bin of its
// low field is (byte & mask_lo), added at entry off_lo + bin of its group's
// histogram unless it equals mfb_lo (the feature's most frequent bin, never
// accumulated); a byte holding two 4-bit fields adds its high field the same
// way. Fields without a feature have mask 0 and mfb 0, so they never add.
struct ByteDev {
  int off_lo, off_hi;
  unsigned mask_lo, mask_hi, mfb_lo, mfb_hi;
};

// Byte p of a packed row as the binning pass writes it: the features binned
// into its low and high bits (-1: none; hi >= 0 only for 4-bit fields).
struct BinByte {
  int lo, hi;
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

// Leaves split per growth round (the tree grows in rounds; see grow.hip).
constexpr int kMaxBatch = 32;

// The tree-growth state, one record per context. Leaves live in slots; the
// tree's own leaf numbering (the reference's, in split order) is the
// position of a slot in the current position array.
struct StepDev {
  int done;          // the tree is complete
  int nseq;          // leaves of the tree (positions)
  int npending;      // leaves split ahead of order, not yet replayed
  int nslots;        // leaf slots in use
  int num_splits;    // splits replayed so far (log entries)
  int parity;        // current position array (0 / 1)
  int search_base;   // first slot searched by the last split pass ...
  int nsearched;     // ... and how many (feat_split rows 0..nsearched)
  int rounds;        // growth rounds run on this tree
  int hist_items;    // histograms built (segments), root included
  long long hist_rows;   // rows summed into histograms, root included
  long long part_rows;   // rows moved by partition passes
  int computed;      // splits computed (replayed or not)
  int pad2;
  double root_output;  // the root's output before any split (learner/mod.rs:647)
  unsigned long long lcur[kMaxBatch], rcur[kMaxBatch];   // rows placed left / right per item
};

// One leaf of the tree in sequential numbering: its slot and best split
// (gain, feature); when the leaf was split ahead of order (spec >= 0), its
// child slots are spec and spec + 1, with their best splits.
struct PosRec {
  double gain;
  double cgain[2];
  int feat, slot, spec;
  int cfeat[2];
  int pad;
};

// A split of the current round.
struct PlanItem {
  SplitRec rec;
  FeatDev fd;               // the split feature
  long long begin, count;   // the parent's segment
  int parent, lslot;        // slots: parent, left child (right = lslot + 1)
  int buf, depth;           // the parent's buffer and depth
  int tile0, pad;           // first partition tile
};

// A split replayed at the start of a round, in sequential order.
struct PlanRep {
  int pos, parent, lslot, pad;
};

// The plan of one round, written by the partition pass's block 0.
struct PlanDev {
  int done;         // the replay completed the tree
  int nwork;        // splits this round
  int total_tiles;
  int nseq, npending, nslots, num_splits;   // state after this round
  int nrep;         // replayed splits (PlanRep entries)
  int complete;     // no child of this round can be split: skip their searches
  int pad;
  PlanItem item[kMaxBatch];
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
  unsigned qmax_g, qmax_h;                  // max |gq| and max u8(hq) of the tree's discretized rows
  // a quantized tree on a full-data bag (quant.rs:144-158 bag_root_sums): the
  // root's sums of f32(gq * f32 grad_scale) and f32(hq * f32 hess_scale) as
  // exact integers in units of 2^bag_eg / 2^bag_eh, and the packed sum read
  // by position 0 .. bag size - 1
  int full_bag;
  int bag_eg, bag_eh;
  int pad2;
  long long bag_sg_fix, bag_sh_fix, bag_isum;
};

}  // namespace lgbmhip
