// Host driver of the MI355X GBDT engine and the C ABI (include/lgbm_hip.h).
//
// The driver mirrors the reference's boosting loop (boosting.rs:957-1139) on
// the host: it launches the device-resident tree-growth sequence for each
// class, copies the split log back once per iteration and replays it into
// Tree arrays (tree.rs:254-334), which stay on the host for the model.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "engine.h"
#include "textnum.h"
#include "metrics.h"
#include "kernels.hip"
#include "grow.hip"

namespace lgbmhip {

#define HIP_CHECK(x)                                                                             \
  do {                                                                                           \
    hipError_t e_ = (x);                                                                         \
    if (e_ != hipSuccess)                                                                        \
      throw Error(LGBMHIP_INTERNAL, std::string(#x) + ": " + hipGetErrorString(e_));             \
  } while (0)


// mt19937.rs:14-88 (the quantizer's generators).
struct Mt {
  uint32_t st[624];
  int i;
  explicit Mt(uint32_t seed) {
    st[0] = seed;
    for (int k = 1; k < 624; ++k) st[k] = 1812433253u * (st[k - 1] ^ (st[k - 1] >> 30)) + (uint32_t)k;
    i = 624;
  }
  uint32_t next() {
    if (i >= 624) {
      for (int k = 0; k < 624; ++k) {
        uint32_t y = (st[k] & 0x80000000u) | (st[(k + 1) % 624] & 0x7fffffffu);
        st[k] = st[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
      }
      i = 0;
    }
    uint32_t z = st[i++];
    z ^= z >> 11;
    z ^= (z << 7) & 0x9d2c5680u;
    z ^= (z << 15) & 0xefc60000u;
    z ^= z >> 18;
    return z;
  }
  double uniform01() {
    const double R = 4294967296.0;
    double s = (double)next();
    s += (double)next() * R;
    double r = s / (R * R);
    return r >= 1.0 ? std::nextafter(1.0, 0.0) : r;
  }
  int uniform_int(int a, int b) {
    uint32_t range = (uint32_t)((int64_t)b - (int64_t)a + 1);
    uint64_t prod = (uint64_t)next() * range;
    uint32_t low = (uint32_t)prod;
    if (low < range) {
      uint32_t thr = (uint32_t)(-range) % range;
      while (low < thr) {
        prod = (uint64_t)next() * range;
        low = (uint32_t)prod;
      }
    }
    return (int)((int64_t)(prod >> 32) + a);
  }
};

// Tree arrays of the model (tree.rs:120-159).
struct HostTree {
  int num_leaves = 1;
  std::vector<int> split_feature_inner, split_feature, left_child, right_child, leaf_parent{-1}, leaf_count{0},
      leaf_depth{0}, internal_count;
  std::vector<float> split_gain;
  std::vector<uint32_t> threshold_in_bin;
  std::vector<double> threshold, leaf_value{0.0}, leaf_weight{0.0}, internal_value, internal_weight;
  std::vector<int8_t> decision_type;
  int num_cat = 0;
  std::vector<int> cat_boundaries{0}, cat_boundaries_inner{0};
  std::vector<uint32_t> cat_threshold, cat_threshold_inner;
  double shrinkage = 1.0;
  static double r0(double v) { return (v >= -(double)1e-35f && v <= (double)1e-35f) ? 0.0 : v; }
  int split(int leaf, int fi, int fr, uint32_t tb, double thr, double lv, double rv, int lc, int rc, double lw,
            double rw, float gain, int mt, bool dl) {
    int node = num_leaves - 1, nl = num_leaves;
    int parent = leaf_parent[leaf];
    if (parent >= 0) {
      if (left_child[parent] == ~leaf) left_child[parent] = node; else right_child[parent] = node;
    }
    split_feature_inner.push_back(fi);
    split_feature.push_back(fr);
    split_gain.push_back(gain);
    left_child.push_back(~leaf);
    right_child.push_back(~nl);
    leaf_parent[leaf] = node;
    leaf_parent.push_back(node);
    internal_weight.push_back(lw + rw);
    internal_value.push_back(leaf_value[leaf]);
    internal_count.push_back(lc + rc);
    leaf_value[leaf] = std::isnan(lv) ? 0.0 : lv;
    leaf_weight[leaf] = lw;
    leaf_count[leaf] = lc;
    leaf_value.push_back(std::isnan(rv) ? 0.0 : rv);
    leaf_weight.push_back(rw);
    leaf_count.push_back(rc);
    int d = leaf_depth[leaf] + 1;
    leaf_depth[leaf] = d;
    leaf_depth.push_back(d);
    int8_t dt = dl ? 2 : 0;
    dt = (int8_t)((dt & 3) | (mt << 2));
    decision_type.push_back(dt);
    threshold_in_bin.push_back(tb);
    threshold.push_back(thr);
    num_leaves++;
    return nl;
  }
  void shrink(double rate) {
    for (int i = 0; i < num_leaves - 1; ++i) {
      leaf_value[i] = r0(leaf_value[i] * rate);
      internal_value[i] = r0(internal_value[i] * rate);
    }
    leaf_value[num_leaves - 1] = r0(leaf_value[num_leaves - 1] * rate);
    shrinkage *= rate;
  }
  void add_bias(double v) {
    for (int i = 0; i < num_leaves - 1; ++i) {
      leaf_value[i] = r0(leaf_value[i] + v);
      internal_value[i] = r0(internal_value[i] + v);
    }
    leaf_value[num_leaves - 1] = r0(leaf_value[num_leaves - 1] + v);
    shrinkage = 1.0;
  }
  void as_constant(double v, int count) {
    *this = HostTree();
    leaf_value[0] = v;
    leaf_count[0] = count;
  }
};

// The header of a model text (io.rs:19-37) and its parameters block.
struct ModelHeader {
  int num_class = 1, ntpi = 1, label_index = 0, ncol = 0;
  bool has_objective = true;                    // false: a custom objective's model (no objective line)
  std::string objective;                        // as written: "binary sigmoid:1", "multiclass num_class:3", ...
  std::string feature_names, feature_infos;     // space-separated, one per column
  bool has_parameters = false;
  std::string parameters;                       // the block's lines, each ending in '\n'
};

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    free();
    if (count == 0) return;
    HIP_CHECK(hipMalloc((void**)&p, count * sizeof(T)));
    n = count;
  }
  void zero(hipStream_t s) {
    if (p) HIP_CHECK(hipMemsetAsync(p, 0, n * sizeof(T), s));
  }
  void free() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { free(); }
};

// The growth state of one tree in flight: its stream, the ordered buffers
// 1-2 and every class's root gradients (buffer 0; buffer 0's bins are the
// context's), the histogram units, the leaf slots and the round's plan. Lane
// l grows the trees of classes l, l + NL, l + 2 NL, ... of an iteration, so
// up to NL trees of one iteration grow at the same time (their gradients are
// all computed before any tree, boosting.rs:988-1001).
struct Lane {
  hipStream_t stream = nullptr;
  bool own_stream = false;
  hipEvent_t done_ev = nullptr;
  DevBuf<uint8_t> gh[3];
  DevBuf<uint8_t> grp[3];      // [0] unused: buffer 0's bins are Ctx::grp0
  DevBuf<int16_t> q_all;       // quantized (gq, hq) of every row (a quantized tree on a full-data bag)
  DevBuf<NodeDev> nodes;       // the lane's last tree's nodes
  DevBuf<int> leaf_shift;      // [L][2] renewal fixed-point shifts
  DevBuf<long long> acc;       // each item's histogram summed over its chunks, one unit per group
  DevBuf<int> acc_units;       // {0, G, 2G, ...}: acc read as one unit per (item, group)
  DevBuf<long long> slab;      // histogram units of one histogram pass
  DevBuf<int> units;           // first unit of each item, total, item count, slab form of each item
  DevBuf<double> pool;
  DevBuf<long long> ipool;
  DevBuf<SplitRec> feat_split, best, log;
  DevBuf<uint8_t> spl;
  DevBuf<StepDev> st;
  DevBuf<PlanDev> plan;
  DevBuf<PlanRep> reps;
  DevBuf<PosRec> pos;
  DevBuf<LeafDev> leaves;
  DevBuf<long long> renew;
  DevBuf<uint16_t> leaf16;     // each row's leaf from the renewal walk (quantized renew)
  ~Lane() {
    if (done_ev) (void)hipEventDestroy(done_ev);
    if (own_stream && stream) (void)hipStreamDestroy(stream);
  }
};

// Trees of one iteration grown at the same time, at most: the device's
// default number of hardware queues per process.
constexpr int kMaxLanes = 16;

// Growth lanes run on concurrent streams: one per hardware queue the HIP
// runtime gives this process (GPU_MAX_HW_QUEUES, 4 when unset), at most 16.
static int hw_queue_lanes() {
  const char* v = std::getenv("GPU_MAX_HW_QUEUES");
  const int q = v ? std::atoi(v) : 4;
  return std::max(1, std::min(kMaxLanes, q > 0 ? q : 4));
}

// A validation set (boosting.rs:378-411): rows binned with the training
// data's mappers, row-major [n][S] like the training bins, and its scores
// [K][n], which every tree's score walk also updates (boosting.rs:1381-1408).
struct ValidData {
  int64_t n = 0;
  DevBuf<uint8_t> bins;
  DevBuf<double> score, init;
  bool has_init = false;
  std::vector<float> label;
  DevBuf<float> d_label, d_weight;
  bool has_weight = false;
  double sum_weights = 0.0;
  std::vector<int32_t> qb;
  DevBuf<int> d_qb;
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int cus = 256;
  size_t lds_max = 65536;
  // data
  int64_t N = 0;
  int F = 0, S = 0, B = 0, ncol = 0;
  // packed rows: W bytes (the u8 features in order, then the features of at
  // most 16 bins two per byte; pack_u4 = 0 keeps every feature a byte)
  int W = 0;
  int pack_u4 = 1;
  bool has_u4 = false;
  std::vector<BinByte> byte_feats;   // [W] the features of each byte
  std::vector<lgbmhip_feature_meta> metas;
  std::vector<std::vector<double>> upper;
  std::vector<FeatDev> feats;
  DevBuf<uint8_t> bins_rm, bins_cm;
  DevBuf<FeatDev> d_feats;
  DevBuf<ByteDev> d_bytes;        // [NG * FB] the histogram kernels' view of each byte (padding included)
  DevBuf<int> d_ginfo;            // [NG][2] each histogram group's entries and whether it holds 4-bit fields
  DevBuf<BinByte> d_binbyte;      // [W] byte_feats on the device (binning into packed rows)
  // value_to_bin tables of the used features (device binning)
  DevBuf<BinFeat> d_binfeat;
  DevBuf<double> d_bounds;
  DevBuf<int> d_ckeys, d_cbins;
  // metadata
  std::vector<float> label;
  DevBuf<float> d_label, d_weight;
  bool has_weight = false;
  double sum_weights = 0.0;   // Metric::new's weight total (metric.rs:238-247)
  std::vector<int32_t> qb;
  // parameters
  SplitParamsDev sp{};
  bool sp_set = false;
  lgbmhip_train_opts opts{};
  bool opts_set = false;
  int K = 1, L = 31;
  // training state
  bool state_ready = false;
  DevBuf<double> score;
  DevBuf<float> g, h;
  // three buffers of (gradients, group-major bins) in leaf order: buffer 0
  // holds the root in row order, leaves alternate between buffers 1 and 2.
  // Buffer 0's bins never change and are shared by the lanes.
  DevBuf<uint8_t> grp0;
  DevBuf<unsigned> rows_buf;   // stage entry points: caller row lists
  std::vector<std::unique_ptr<Lane>> lanes;
  int NL = 1;
  hipEvent_t grad_ev = nullptr;
  // growth state: leaf slots (MS of them) per lane
  int MS = 0;
  DevBuf<IterDev> it;
  DevBuf<double> leafval, leafraw;
  DevBuf<double> grv, hrv;
  DevBuf<int> starts;
  DevBuf<int> call_ctr;
  DevBuf<int8_t> gq, hq;
  std::unique_ptr<Mt> start_eng;
  int n_calls = 0;
  int hist_tpr = 8, hist_fb = 0, hist_ng = 1, hist_chunks = 256, hist_min_rows = 2048;
  int hist_pu = 0;       // entries of a histogram unit (the largest group's)
  // growth rounds from this index on search straight from the histogram
  // units (no slab reduction launch); 0: never; -1 in tune_late_round: the
  // default
  int late_round = 12, tune_late_round = -1;
  int num_groups = 0;   // the reference's feature-group count (0: one group per used feature)
  int hist_grid = 256;   // blocks of a histogram launch (they loop over the units)
  size_t hist_lds = 0;
  int part_grid = 1024;
  int tune_part_grid = 0, tune_hist_chunks = 0, tune_hist_min_rows = 0, tune_hist_lds = 0;
  // pinned copies of the per-tree results
  std::vector<SplitRec*> h_log;
  std::vector<double*> h_raw;
  std::vector<StepDev*> h_st;
  // graphs: per class, one tree (init, root, rounds_main rounds, tail) and
  // a continuation (rounds_cont rounds, tail) for trees that need more rounds
  // per class: captured main launches and continuations, by (rounds, score update, learning rate)
  struct GraphKey {
    int rounds;
    bool score_update;
    double lr;
    bool operator==(const GraphKey& o) const {
      return rounds == o.rounds && score_update == o.score_update && lr == o.lr;
    }
  };
  std::vector<std::vector<std::pair<GraphKey, hipGraphExec_t>>> graphs, graphs_cont;
  // batched trees: the class trees of an iteration share growth launches,
  // `batch` trees per launch (lane k holds class k's tree; batch g, classes
  // g * batch .. , runs on its first lane's stream), each kernel finding its
  // tree's arguments in these tables (one entry per class)
  int batch = 0;
  int nbatch = 0;
  DevBuf<PartArgs> tab_part;
  DevBuf<HistArgs> tab_hist_root, tab_hist_round;
  DevBuf<SplitArgs> tab_split_root, tab_split_round, tab_split_raw;
  DevBuf<SlabTab> tab_slab;
  std::vector<PartArgs> h_tab_part;
  std::vector<HistArgs> h_tab_hist_root, h_tab_hist_round;
  std::vector<SplitArgs> h_tab_split_root, h_tab_split_round, h_tab_split_raw;
  std::vector<SlabTab> h_tab_slab;
  std::vector<std::vector<std::pair<GraphKey, hipGraphExec_t>>> graphs_batch;   // [2 g] main, [2 g + 1] continuation
  std::vector<int> want_rounds;   // rounds the class's previous tree ran to completion
  int rounds_main = 24, rounds_cont = 8;
  std::vector<int> last_rounds;   // rounds launched for each class's last tree
  // model
  std::vector<HostTree> trees;
  bool models_empty_init = true;
  int iter = 0;
  // lambdarank state (rank.rs:129-341)
  DevBuf<int> d_qb, r_rank, r_order;
  DevBuf<double> r_imd, r_table, r_gain, r_disc;
  DevBuf<float> r_lam, r_hes, r_bl;
  DevBuf<double> r_bs;
  double r_min_in = 0, r_max_in = 0, r_factor = 0;
  // row bagging (sample_strategy.rs:186-318): the bag drawn on the device,
  // the root's rows of every tree until the next draw, and buffer 0's bins
  // for it (bag_grp0, group-major, positions 0 .. bag_m - 1)
  bool bag_on = false, bag_subset = false, need_re_bagging = false, bag_balanced = false;
  bool bag_active = false;   // the current trees grow on the bag
  bool goss_on = false;
  DevBuf<GossBlock> goss_blocks;
  DevBuf<float> goss_abs, goss_thr;
  int goss_nblock = 0;
  int64_t bag_cnt = 0, bag_m = 0;
  DevBuf<unsigned> bag_rows, bag_lcg;
  DevBuf<uint8_t> bag_flags, bag_grp0;
  bool bag_by_query = false;   // bagging_by_query: whole queries drawn
  DevBuf<int> bag_qb;          // their boundaries on the device
  DevBuf<int> bag_counts, bag_offsets;
  DevBuf<long long> bag_total;
  // validation sets and the metric workspace
  std::vector<std::unique_ptr<ValidData>> valid;
  MetricWs* mws = nullptr;
  DevBuf<int> d_qb_eval;   // training query boundaries for NDCG when the objective keeps none
  bool loaded_model = false;   // trees from model text: raw prediction only
  ModelHeader loaded_head;     // the loaded text's header and parameters block
  // prediction window and early stopping (predict.rs:178-202): iterations
  // start .. start + num (num <= 0: to the end), margin test every freq (0: off)
  int pred_start = 0, pred_num = 0, pred_es_freq = 0;
  double pred_es_margin = 0.0;
  // prediction cache
  bool pred_dirty = true;
  DevBuf<uint4> p_nodes;
  DevBuf<int> p_node_off, p_leaf_off, p_nleaves;
  DevBuf<double> p_leafval;
  DevBuf<unsigned> p_cat;
  int p_T = 0;
  bool p_binned = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  // profiling: HIP events around every histogram / split / partition launch
  int profile = 0;
  std::vector<hipEvent_t> pev;   // [K][L + 2 steps][4 kinds][2]
  double prof[12] = {0};         // hist_ms, hist_launches, hist_rows, hist_bytes, split_ms, part_ms, iter_ms, iters,
                                 // rows partitioned, splits computed, splits used

  ~Ctx() {
    lanes.clear();
    if (mws) metric_ws_free(mws);
    if (grad_ev) (void)hipEventDestroy(grad_ev);
    for (auto e : pev) if (e) (void)hipEventDestroy(e);
    for (auto* gs : {&graphs, &graphs_cont, &graphs_batch})
      for (auto& v : *gs)
        for (auto& g : v) (void)hipGraphExecDestroy(g.second);
    for (auto p : h_log) if (p) (void)hipHostFree(p);
    for (auto p : h_raw) if (p) (void)hipHostFree(p);
    for (auto p : h_st) if (p) (void)hipHostFree(p);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

static void set_device(Ctx* c) { HIP_CHECK(hipSetDevice(c->device)); }

// The lane that grows class k's trees.
static Lane& lane_of(Ctx* c, int k) { return *c->lanes[(size_t)(k % c->NL)]; }

static int grid_for(int64_t n, int threads, int cap) {
  int64_t g = (n + threads - 1) / threads;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, cap));
}

// Buffer-0 bins of rows[0..m) (every row when rows is null) into the
// group-major layout of the ordered buffers (row stride N), on stream s.
static void launch_group_rows(Ctx* c, const unsigned* rows, int64_t m, uint8_t* out, size_t out_bytes,
                              hipStream_t s) {
  if (m <= 0) return;
  // the kernels write groups [NG][N][FB] at row stride N: the buffer must hold them
  if ((size_t)c->hist_ng * (size_t)c->N * (size_t)c->hist_fb > out_bytes)
    throw Error(LGBMHIP_INTERNAL, "group-major row buffer smaller than the feature groups");
  const int CH = c->hist_ng * c->hist_fb / 16;
  const bool vec = c->S % 16 == 0 && c->hist_fb % 16 == 0 && c->hist_ng * c->hist_fb <= c->S && CH >= 1 && CH <= 8;
  if (vec) {
    const dim3 g(grid_for(m, 256, c->cus * 16)), b(256);
    switch (CH) {
      case 1: hipLaunchKernelGGL(k_group_rows_v<1>, g, b, 0, s, rows, (long long)m, c->bins_rm.p, c->S, c->hist_fb, (long long)c->N, out); break;
      case 2: hipLaunchKernelGGL(k_group_rows_v<2>, g, b, 0, s, rows, (long long)m, c->bins_rm.p, c->S, c->hist_fb, (long long)c->N, out); break;
      case 3: hipLaunchKernelGGL(k_group_rows_v<3>, g, b, 0, s, rows, (long long)m, c->bins_rm.p, c->S, c->hist_fb, (long long)c->N, out); break;
      case 4: hipLaunchKernelGGL(k_group_rows_v<4>, g, b, 0, s, rows, (long long)m, c->bins_rm.p, c->S, c->hist_fb, (long long)c->N, out); break;
      case 5: hipLaunchKernelGGL(k_group_rows_v<5>, g, b, 0, s, rows, (long long)m, c->bins_rm.p, c->S, c->hist_fb, (long long)c->N, out); break;
      case 6: hipLaunchKernelGGL(k_group_rows_v<6>, g, b, 0, s, rows, (long long)m, c->bins_rm.p, c->S, c->hist_fb, (long long)c->N, out); break;
      case 7: hipLaunchKernelGGL(k_group_rows_v<7>, g, b, 0, s, rows, (long long)m, c->bins_rm.p, c->S, c->hist_fb, (long long)c->N, out); break;
      default: hipLaunchKernelGGL(k_group_rows_v<8>, g, b, 0, s, rows, (long long)m, c->bins_rm.p, c->S, c->hist_fb, (long long)c->N, out); break;
    }
  } else {
    hipLaunchKernelGGL(k_group_rows, dim3(grid_for(m * c->hist_ng * (c->hist_fb / 4), 256, c->cus * 16)), dim3(256), 0,
                       s, rows, (long long)m, c->bins_rm.p, c->S, c->W, c->hist_fb, c->hist_ng, (long long)c->N, out);
  }
  HIP_CHECK(hipGetLastError());
}

// Row-major stride: rows of <= 32 bytes padded to a power of two (one 32-byte
// sector per row gather), wider rows to 16 bytes.
static int row_stride(int F) {
  if (F <= 32) {
    int s = 4;
    while (s < F) s <<= 1;
    return s;
  }
  return (F + 15) / 16 * 16;
}

// The used features of the training data: metas, bin bounds, split-search
// records and the device tables of value_to_bin (bounds of numerical
// features; categories of categorical ones, sorted, with their bins).
static void set_features(Ctx* c, const lgbmhip_feature_meta* metas, int F, const double* upper, int64_t N, int ncol) {
  if (F <= 0 || N <= 0) throw Error(LGBMHIP_INVALID_DATA, "no rows or no used features");
  if (N > INT32_MAX) throw Error(LGBMHIP_INVALID_DATA, "more than 2^31-1 rows");
  set_device(c);
  c->N = N;
  c->F = F;
  c->ncol = ncol;
  c->metas.assign(metas, metas + F);
  c->upper.clear();
  c->feats.clear();
  c->B = 1;
  size_t off = 0;
  for (int f = 0; f < F; ++f) {
    const auto& m = metas[f];
    if (m.num_bin < 1 || m.num_bin > kMaxBins) throw Error(LGBMHIP_INVALID_DATA, "num_bin out of range [1, 256]");
    if (m.most_freq_bin < 0 || m.most_freq_bin >= m.num_bin || m.default_bin < 0 || m.default_bin >= m.num_bin)
      throw Error(LGBMHIP_INVALID_DATA, "bin index out of range in feature meta");
    c->upper.emplace_back(upper + off, upper + off + m.num_bin);
    off += (size_t)m.num_bin;
    FeatDev d{m.num_bin, m.missing_type, m.default_bin, m.most_freq_bin, m.bin_type, m.most_freq_bin == 0 ? 1 : 0,
              0, -1, 0, 0};
    c->feats.push_back(d);
    c->B = std::max(c->B, m.num_bin);
  }
  // packed rows: bytes of the u8 features in feature order, then the
  // features of at most 16 bins as 4-bit fields, two per byte (Dense4 of
  // bin.rs:92-121 and BinKind::dense :314-324 pick 4 bits the same way)
  {
    int n8 = 0, n4 = 0;
    auto u4 = [&](int f) { return c->pack_u4 && c->feats[(size_t)f].num_bin <= 16; };
    for (int f = 0; f < F; ++f) (u4(f) ? n4 : n8)++;
    c->W = n8 + (n4 + 1) / 2;
    c->has_u4 = n4 > 0;
    c->byte_feats.assign((size_t)c->W, BinByte{-1, -1});
    int p8 = 0, j4 = 0;
    for (int f = 0; f < F; ++f) {
      FeatDev& d = c->feats[(size_t)f];
      if (u4(f)) {
        d.pos = n8 + j4 / 2;
        d.nib = j4 % 2;
        (d.nib ? c->byte_feats[(size_t)d.pos].hi : c->byte_feats[(size_t)d.pos].lo) = f;
        ++j4;
      } else {
        d.pos = p8++;
        d.nib = -1;
        c->byte_feats[(size_t)d.pos].lo = f;
      }
    }
  }
  c->S = row_stride(c->W);
  std::vector<BinFeat> bfs;
  std::vector<double> bounds;
  std::vector<int> ck, cb;
  for (int f = 0; f < F; ++f) {
    const auto& m = metas[f];
    if (m.real_index < 0 || m.real_index >= ncol) throw Error(LGBMHIP_INVALID_DATA, "feature column out of range");
    BinFeat bf{m.real_index, m.num_bin, m.missing_type, m.bin_type, 0, 0};
    if (m.bin_type == 1) {
      std::vector<std::pair<int, int>> cats;
      for (int b = 0; b < m.num_bin; ++b)
        if (c->upper[f][b] >= 0.0) cats.emplace_back(sat_i32(c->upper[f][b]), b);
      std::sort(cats.begin(), cats.end());
      bf.off = (int)ck.size();
      bf.ncat = (int)cats.size();
      for (auto& pr : cats) {
        ck.push_back(pr.first);
        cb.push_back(pr.second);
      }
    } else {
      bf.off = (int)bounds.size();
      bounds.insert(bounds.end(), c->upper[f].begin(), c->upper[f].end());
    }
    bfs.push_back(bf);
  }
  c->d_binfeat.alloc(F);
  c->d_bounds.alloc(std::max<size_t>(bounds.size(), 1));
  c->d_ckeys.alloc(std::max<size_t>(ck.size(), 1));
  c->d_cbins.alloc(std::max<size_t>(cb.size(), 1));
  HIP_CHECK(hipMemcpy(c->d_binfeat.p, bfs.data(), F * sizeof(BinFeat), hipMemcpyHostToDevice));
  if (!bounds.empty()) HIP_CHECK(hipMemcpy(c->d_bounds.p, bounds.data(), bounds.size() * 8, hipMemcpyHostToDevice));
  if (!ck.empty()) {
    HIP_CHECK(hipMemcpy(c->d_ckeys.p, ck.data(), ck.size() * 4, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(c->d_cbins.p, cb.data(), cb.size() * 4, hipMemcpyHostToDevice));
  }
  c->d_feats.alloc(F);
  HIP_CHECK(hipMemcpy(c->d_feats.p, c->feats.data(), F * sizeof(FeatDev), hipMemcpyHostToDevice));
  c->d_binbyte.alloc((size_t)c->W);
  HIP_CHECK(hipMemcpy(c->d_binbyte.p, c->byte_feats.data(), c->W * sizeof(BinByte), hipMemcpyHostToDevice));
  c->state_ready = false;
  c->trees.clear();
  c->pred_dirty = true;
  c->valid.clear();
}

// One row of u8 bins (F of them, feature order) into its packed form (out
// holds zeros: the row's S bytes).
static void pack_row(const Ctx* c, const uint8_t* in, uint8_t* out) {
  for (int f = 0; f < c->F; ++f) {
    const FeatDev& d = c->feats[(size_t)f];
    if (d.nib < 0) out[d.pos] = in[f];
    else out[d.pos] = (uint8_t)(out[d.pos] | ((in[f] & 0xF) << (4 * d.nib)));
  }
}

static void upload_bins(Ctx* c, const lgbmhip_feature_meta* metas, int F, const double* upper, const uint8_t* bins,
                        int64_t N, int ncol) {
  set_features(c, metas, F, upper, N, ncol);
  for (int64_t i = 0; i < N * F; ++i)
    if (bins[i] >= (uint8_t)c->metas[i % F].num_bin && c->metas[i % F].num_bin < 256)
      throw Error(LGBMHIP_INVALID_DATA, "bin value >= num_bin");
  std::vector<uint8_t> rm((size_t)N * c->S, 0), cm((size_t)N * F);
  for (int64_t r = 0; r < N; ++r) {
    pack_row(c, bins + (size_t)r * F, &rm[(size_t)r * c->S]);
    for (int f = 0; f < F; ++f) cm[(size_t)f * N + r] = bins[(size_t)r * F + f];
  }
  c->bins_rm.alloc(rm.size());
  c->bins_cm.alloc(cm.size());
  HIP_CHECK(hipMemcpy(c->bins_rm.p, rm.data(), rm.size(), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(c->bins_cm.p, cm.data(), cm.size(), hipMemcpyHostToDevice));
}

// Bins device rows (rows x ncol f32) with the training mappers into rows of
// `stride` bytes on the context's stream: packed rows (the training layout)
// or, with packed = false, one u8 bin per feature.
static void launch_bin_rows(Ctx* c, const float* dX, int64_t rows, int ncol, uint8_t* dout, int stride,
                            bool packed) {
  if (ncol != c->ncol) throw Error(LGBMHIP_INVALID_DATA, "column count differs from the training data");
  if (rows <= 0) return;
  const int64_t work = rows * ((stride + 3) / 4);
  const int nbounds = (int)c->d_bounds.n;
  const size_t lds = (size_t)nbounds * 8;
  const BinByte* map = c->d_binbyte.p;
  if (lds <= 64 * 1024) {
    // every feature's bounds in LDS: 1024-thread blocks, two per CU at 64 KiB
    auto k = packed ? k_bin_rows<true, true> : k_bin_rows<true, false>;
    HIP_CHECK(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k, dim3(grid_for(work, 1024, c->cus * 4)), dim3(1024), lds, c->stream, dX, (long long)rows,
                       ncol, c->d_binfeat.p, c->F, c->d_bounds.p, nbounds, c->d_ckeys.p, c->d_cbins.p, dout, stride,
                       map, c->W);
  } else {
    auto k = packed ? k_bin_rows<false, true> : k_bin_rows<false, false>;
    hipLaunchKernelGGL(k, dim3(grid_for(work, 256, c->cus * 16)), dim3(256), 0, c->stream, dX, (long long)rows, ncol,
                       c->d_binfeat.p, c->F, c->d_bounds.p, nbounds, c->d_ckeys.p, c->d_cbins.p, dout, stride, map,
                       c->W);
  }
  HIP_CHECK(hipGetLastError());
}

// Dataset::from_matrix's binning pass (dataset.rs:355-422, value_to_bin
// binning.rs:637-667) on the device: the raw rows (host memory, staged in
// chunks, or device memory) binned straight into the row-major bins, then
// the column-major copy.
static void upload_raw(Ctx* c, const Binned& b, const float* X, int64_t N, int ncol, bool on_device) {
  if (ncol != b.ncol) throw Error(LGBMHIP_INVALID_DATA, "column count differs from the mappers");
  const int F = (int)b.used.size();
  std::vector<lgbmhip_feature_meta> metas(F);
  std::vector<double> upper;
  for (int f = 0; f < F; ++f) {
    const Mapper& m = b.mappers[b.used[f]];
    metas[f] = lgbmhip_feature_meta{b.used[f], m.num_bin, m.missing_type, (int32_t)m.default_bin,
                                    (int32_t)m.most_freq_bin, m.bin_type};
    for (int i = 0; i < m.num_bin; ++i) upper.push_back(m.bin_type == 1 ? (double)m.categories[i] : m.bounds[i]);
  }
  set_features(c, metas.data(), F, upper.data(), N, ncol);
  c->bins_rm.alloc((size_t)N * c->S);
  c->bins_cm.alloc((size_t)N * F);
  if (on_device) {
    launch_bin_rows(c, X, N, ncol, c->bins_rm.p, c->S, true);
  } else {
    const int64_t chunk = std::max<int64_t>(1, ((int64_t)64 << 20) / ncol);   // 256 MiB of f32 per copy
    DevBuf<float> stage;
    stage.alloc((size_t)std::min(N, chunk) * ncol);
    for (int64_t r0 = 0; r0 < N; r0 += chunk) {
      const int64_t m = std::min(chunk, N - r0);
      HIP_CHECK(hipMemcpyAsync(stage.p, X + (size_t)r0 * ncol, (size_t)m * ncol * 4, hipMemcpyHostToDevice, c->stream));
      launch_bin_rows(c, stage.p, m, ncol, c->bins_rm.p + (size_t)r0 * c->S, c->S, true);
      HIP_CHECK(hipStreamSynchronize(c->stream));
    }
  }
  hipLaunchKernelGGL(k_rm_to_cm, dim3(grid_for(N * F, 256, c->cus * 16)), dim3(256), 0, c->stream, c->bins_rm.p,
                     (long long)N, c->S, F, c->d_feats.p, c->bins_cm.p);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(c->stream));
}

// Histogram launch shape over the packed row's bytes: bytes per group so the
// block's LDS histogram fits the budget (a u8 feature's byte takes B entries,
// a byte of two 4-bit fields 32); threads per row = 4-byte chunks of the
// group, to a power of 2. Then each feature's group and first entry in its
// group's histogram unit, each byte's view for the histogram kernels, and
// each group's entry count (a unit holds hist_pu entries, the largest).
static void plan_hist(Ctx* c, bool quant) {
  const int W = c->W, B = c->B;
  const size_t per_feat = (size_t)B * 2 * (quant ? 4 : 8);
  // 32 KiB by default: five blocks share a CU's LDS and each block zeroes and
  // flushes a small histogram, which wins over wide feature groups at every
  // leaf size measured on 10.5M x 28 (quantized and FP64 forms)
  size_t budget = std::min<size_t>(c->lds_max, 160 * 1024);
  budget = std::min<size_t>(budget, c->tune_hist_lds > 0 ? (size_t)c->tune_hist_lds : (size_t)32 * 1024);
  // at most 128 bytes (32 threads of 4 bins) of a row per block
  int fb = (int)std::min<size_t>(std::min<size_t>((size_t)W, 128), budget / per_feat);
  fb = std::max(4, fb / 4 * 4);
  // the quantized form also holds 32-bit bins and trash words beside its
  // int64 sums (about 12 bytes per bin): narrow the group until that fits
  if (quant)
    while (fb > 4 && q32_lds_bytes(std::min(fb, W) * B) > std::min<size_t>(c->lds_max, 160 * 1024)) fb -= 4;
  if (fb >= W) fb = (W + 3) / 4 * 4;
  int ng = (W + fb - 1) / fb;
  // balance the groups
  fb = ((W + ng - 1) / ng + 3) / 4 * 4;
  ng = (W + fb - 1) / fb;
  int chunks4 = fb / 4;
  int tpr = 1;
  while (tpr < chunks4) tpr <<= 1;
  if (tpr > 32) throw Error(LGBMHIP_UNSUPPORTED, "feature group wider than 128 bytes");
  // entries of each group's histogram, in byte order (a 4-bit pair: the low
  // field's 16 bins, then the high field's)
  std::vector<ByteDev> bytes((size_t)ng * fb, ByteDev{0, 0, 0u, 0u, 0u, 0u});
  std::vector<int> ginfo((size_t)2 * ng, 0);   // per group: entries, holds 4-bit fields
  int pu = 0;
  for (int g = 0; g < ng; ++g) {
    int off = 0;
    for (int p = g * fb; p < std::min(W, (g + 1) * fb); ++p) {
      const BinByte bf = c->byte_feats[(size_t)p];
      ByteDev& bd = bytes[(size_t)p];
      FeatDev& lo = c->feats[(size_t)bf.lo];
      lo.hg = g;
      lo.hoff = off;
      bd.off_lo = off;
      bd.mfb_lo = (unsigned)lo.mfb;
      if (lo.nib < 0) {
        bd.mask_lo = 0xFFu;
        off += B;
      } else {
        bd.mask_lo = 0xFu;
        if (bf.hi >= 0) {
          FeatDev& hi = c->feats[(size_t)bf.hi];
          hi.hg = g;
          hi.hoff = off + 16;
          bd.off_hi = off + 16;
          bd.mask_hi = 0xFu;
          bd.mfb_hi = (unsigned)hi.mfb;
        }
        off += 32;
        ginfo[(size_t)2 * g + 1] = 1;
      }
    }
    ginfo[(size_t)2 * g] = off;
    pu = std::max(pu, off);
  }
  c->hist_fb = fb;
  c->hist_ng = ng;
  c->hist_tpr = tpr;
  c->hist_pu = pu;
  c->hist_lds = (size_t)pu * (quant ? 8 : 16);
  // the quantized form's 32-bit bins, trash words and int64 sums
  if (quant) c->hist_lds = std::max(c->hist_lds, q32_lds_bytes(pu));
  if (c->hist_lds > c->lds_max) throw Error(LGBMHIP_UNSUPPORTED, "histogram does not fit in LDS");
  set_hist_lds_limits(c->hist_lds);
  c->d_bytes.alloc(bytes.size());
  HIP_CHECK(hipMemcpy(c->d_bytes.p, bytes.data(), bytes.size() * sizeof(ByteDev), hipMemcpyHostToDevice));
  c->d_ginfo.alloc(ginfo.size());
  HIP_CHECK(hipMemcpy(c->d_ginfo.p, ginfo.data(), ginfo.size() * sizeof(int), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(c->d_feats.p, c->feats.data(), c->F * sizeof(FeatDev), hipMemcpyHostToDevice));
}

// Host-side tables of the lambdarank objective, computed as the reference
// does (rank.rs:313-328 inverse max DCG, :166-179 sigmoid table, dcg.rs:27-52).
static void setup_rank(Ctx* c) {
  const auto& qb = c->qb;
  if (qb.size() < 2) throw Error(LGBMHIP_INVALID_DATA, "Ranking tasks require query information");
  if (c->has_weight) throw Error(LGBMHIP_UNSUPPORTED, "lambdarank with row weights");
  if (qb.front() != 0 || qb.back() != (int32_t)c->N) throw Error(LGBMHIP_INVALID_DATA, "query boundaries do not cover the data");
  const int Q = (int)qb.size() - 1;
  std::vector<double> gain(31), disc(10000);
  gain[0] = 0.0;
  for (int i = 1; i < 31; ++i) gain[i] = (double)((1 << i) - 1);
  for (int i = 0; i < 10000; ++i) disc[i] = 1.0 / std::log2(2.0 + (double)i);
  for (float l : c->label) {
    if (std::fabs(l - (float)(int)l) > (double)1e-15f) throw Error(LGBMHIP_INVALID_DATA, "label should be int type for ranking task");
    if (l < 0.0f) throw Error(LGBMHIP_INVALID_DATA, "Label should be non-negative for ranking task");
    if ((int)l >= 31) throw Error(LGBMHIP_INVALID_DATA, "Label is not less than the number of label mappings (31)");
  }
  const int k = std::max(c->opts.lambdarank_truncation_level, 0);
  std::vector<double> imd(Q);
  for (int q = 0; q < Q; ++q) {
    const int cnt = qb[q + 1] - qb[q];
    if (cnt < 0) throw Error(LGBMHIP_INVALID_DATA, "query boundaries must ascend");
    if (cnt > 10000) throw Error(LGBMHIP_INVALID_DATA, "Number of rows exceeds upper limit of 10000 for a query");
    std::vector<int> cntl(31, 0);
    for (int i = qb[q]; i < qb[q + 1]; ++i) cntl[(int)c->label[i]]++;
    int top = 30;
    double ret = 0.0;
    for (int j = 0; j < std::min(k, cnt); ++j) {
      while (top > 0 && cntl[top] <= 0) top--;
      if (top < 0) break;
      ret += disc[j] * gain[top];
      cntl[top]--;
    }
    imd[q] = ret > 0.0 ? 1.0 / ret : ret;
  }
  const int bins = 1024 * 1024;
  const double sig = c->opts.sigmoid;
  c->r_min_in = -50.0 / sig / 2.0;
  c->r_max_in = -c->r_min_in;
  c->r_factor = (double)bins / (c->r_max_in - c->r_min_in);
  std::vector<double> table(bins);
  for (int i = 0; i < bins; ++i) {
    const double sc = (double)i / c->r_factor + c->r_min_in;
    table[i] = 1.0 / (1.0 + std::exp(sc * sig));
  }
  c->d_qb.alloc(qb.size());
  c->r_imd.alloc(Q);
  c->r_table.alloc(bins);
  c->r_gain.alloc(31);
  c->r_disc.alloc(10000);
  c->r_rank.alloc(c->N);
  c->r_order.alloc(Q);
  c->r_lam.alloc(c->N);
  c->r_hes.alloc(c->N);
  c->r_bs.alloc(c->N);
  c->r_bl.alloc(c->N);
  HIP_CHECK(hipMemcpy(c->d_qb.p, qb.data(), qb.size() * 4, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(c->r_imd.p, imd.data(), Q * 8, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(c->r_table.p, table.data(), bins * 8, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(c->r_gain.p, gain.data(), 31 * 8, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(c->r_disc.p, disc.data(), 10000 * 8, hipMemcpyHostToDevice));
  // blocks take the longest queries first, so they do not trail the launch
  std::vector<int> order(Q);
  for (int q = 0; q < Q; ++q) order[q] = q;
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return qb[x + 1] - qb[x] > qb[y + 1] - qb[y]; });
  HIP_CHECK(hipMemcpy(c->r_order.p, order.data(), Q * 4, hipMemcpyHostToDevice));
}

// A validation set's scores back to their start: its init scores or zeros
// (boosting.rs:389-396).
static void reset_valid_scores(Ctx* c, ValidData& v) {
  const size_t cnt = (size_t)c->K * v.n;
  if (v.has_init && v.init.n != cnt)
    throw Error(LGBMHIP_INVALID_DATA, "validation init_score size does not match num_class x rows");
  if (v.score.n != cnt) v.score.alloc(cnt);
  if (v.has_init) HIP_CHECK(hipMemcpyAsync(v.score.p, v.init.p, cnt * 8, hipMemcpyDeviceToDevice, c->stream));
  else v.score.zero(c->stream);
}

// Destroys every captured tree launch (their kernels' arguments or the
// validation sets they update changed).
static void drop_graphs(Ctx* c) {
  for (auto* gs : {&c->graphs, &c->graphs_cont, &c->graphs_batch})
    for (auto& v : *gs) {
      for (auto& g : v) (void)hipGraphExecDestroy(g.second);
      v.clear();
    }
}

// Launch shapes of the tree-growth kernels (the tuning fields, else the
// defaults) and the buffers whose layout depends on them: the histogram
// feature groups, units and slabs, and the lanes' group-major row buffers.
// Buffer 0 (the root's rows) is gathered by the caller.
static void set_shapes(Ctx* c, bool quant) {
  const int64_t N = c->N;
  const int B = c->B;
  // defaults from sweeps on 10.5M x 28 (quantized and FP64 forms)
  c->part_grid = (int)std::min<int64_t>((N + part_tile() - 1) / part_tile(),
                                        c->tune_part_grid > 0 ? c->tune_part_grid : 4 * c->cus);
  c->hist_chunks = c->tune_hist_chunks > 0 ? c->tune_hist_chunks : (quant ? 2 : 1) * c->cus;
  // quantized: chunks no longer than an epoch of the 32-bit histogram bins
  // for the quantizer's range (|gq| <= bins / 2, hq <= bins): 4095 rows for
  // 16 bins, so a chunk's bins go to its slab without an int64 fold
  int rows_q = 4096;
  if (quant) {
    const int qb = std::max(2, c->opts.num_grad_quant_bins);
    const int epoch = std::min(32767 / std::max(1, qb / 2), 65535 / qb);
    if (epoch >= 2048) rows_q = std::min(rows_q, epoch);
  }
  c->hist_min_rows = c->tune_hist_min_rows > 0 ? c->tune_hist_min_rows : rows_q;
  // round 12 on 10.5M rows (its leaves take one or two row chunks; measured
  // against rounds 1-32), two rounds later per doubling of rows past 16M
  if (c->tune_late_round >= 0) {
    c->late_round = c->tune_late_round;
  } else {
    c->late_round = 12;
    for (int64_t n = N; n > ((int64_t)16 << 20); n /= 2) c->late_round += 2;
  }
  plan_hist(c, quant);
  // one resident wave of blocks: blocks loop over the units, and blocks
  // beyond what the CUs hold at once only add dispatch latency to the
  // small launches late in a tree
  {
    const int occ = hist_occupancy(quant, c->has_u4, c->hist_tpr, c->hist_lds);
    const int units = c->hist_chunks * c->hist_ng;
    c->hist_grid = occ > 0 ? std::min(units, occ * c->cus) : units;
  }
  // histogram kernels address a chunk's rows with 32-bit byte offsets. A leaf
  // of cnt rows takes min(hist_chunks, ceil(cnt / hist_min_rows)) chunks, so
  // no chunk is longer than max(hist_min_rows, ceil(N / hist_chunks)) rows.
  const int64_t chunk_max = std::max<int64_t>(c->hist_min_rows, (N + c->hist_chunks - 1) / c->hist_chunks);
  if (chunk_max * (int64_t)c->hist_fb >= ((int64_t)1 << 31))
    throw Error(LGBMHIP_UNSUPPORTED,
                "histogram chunks too long for 32-bit row offsets: raise hist_chunks or lower hist_min_rows");
  std::vector<int> au(kMaxBatch + 1);
  for (int w = 0; w <= kMaxBatch; ++w) au[w] = w * c->hist_ng;
  // histogram units of one pass: at most (rows / chunk + splits) per group
  const int64_t max_units = ((N + c->hist_min_rows - 1) / c->hist_min_rows + kMaxBatch + 1) * c->hist_ng;
  const size_t gbytes = (size_t)c->hist_ng * N * c->hist_fb;
  for (auto& lp : c->lanes) {
    Lane& ln = *lp;
    ln.acc.alloc((size_t)kMaxBatch * c->hist_ng * c->hist_pu * (quant ? 1 : 2));
    HIP_CHECK(hipMemcpy(ln.acc_units.p, au.data(), au.size() * sizeof(int), hipMemcpyHostToDevice));
    ln.slab.alloc((size_t)max_units * c->hist_pu * (quant ? 1 : 2));
    for (int b = 1; b < 3; ++b) ln.grp[b].alloc(gbytes);
  }
  // buffer 0 (the root's rows, gathered by the caller) and a bag's buffer 0
  // in the same group layout
  if (c->grp0.n != gbytes) c->grp0.alloc(gbytes);
  if (c->bag_grp0.n && c->bag_grp0.n != gbytes) c->bag_grp0.alloc(gbytes);
  // captured trees hold the previous shapes' launches
  drop_graphs(c);
}

static void setup_training(Ctx* c) {
  if (!c->N) throw Error(LGBMHIP_INVALID_DATA, "no bins uploaded");
  if (c->label.empty()) throw Error(LGBMHIP_INVALID_DATA, "no labels set");
  if (!c->sp_set || !c->opts_set) throw Error(LGBMHIP_INVALID_PARAMETER, "split params / train opts not set");
  set_device(c);
  const auto& o = c->opts;
  const bool quant = o.use_quantized_grad != 0;
  c->K = o.objective == 1 ? o.num_class : 1;
  if (c->K < 1 || c->K > 16) throw Error(LGBMHIP_UNSUPPORTED, "num_class must be in [1, 16]");
  if (o.objective == 2) setup_rank(c);
  c->L = std::max(c->sp.num_leaves, 2);
  if (c->L > 1024) throw Error(LGBMHIP_UNSUPPORTED, "num_leaves above 1024");
  if (c->F >= 65535) throw Error(LGBMHIP_UNSUPPORTED, "65535 or more used features");
  if (quant) {
    if (o.num_grad_quant_bins < 2 || o.num_grad_quant_bins > 126)
      throw Error(LGBMHIP_INVALID_PARAMETER, "num_grad_quant_bins out of range");
  }
  const int64_t N = c->N;
  const int F = c->F, B = c->B, L = c->L, K = c->K;
  c->score.alloc((size_t)K * N);
  c->score.zero(c->stream);
  c->g.alloc((size_t)K * N);
  c->h.alloc((size_t)K * N);
  set_walk_lds_limit(L);
  // leaf slots: at most L - 1 replayed splits plus at most L splits made
  // ahead of order and never replayed, two slots each, plus the root
  c->MS = 4 * L;
  const int MS = c->MS;
  // lanes: lane 0 runs on the context's stream; more lanes get streams of
  // their own (default priority). Batched trees: one lane per class, all on
  // the context's stream.
  // batch_trees 1 picks the size: three batches (on the multiclass
  // benchmark, 10 classes, batches of 4 measured 52 it/s against 48 for 5,
  // 48 for 3, 39 for one batch of 10 and 44 for one tree per launch)
  const int bt = o.batch_trees == 1 ? (K + 2) / 3 : o.batch_trees;
  c->batch = K > 1 && bt > 1 ? std::min(K, bt) : 0;
  c->nbatch = c->batch ? (K + c->batch - 1) / c->batch : 0;
  c->NL = c->batch ? K : std::min(K, hw_queue_lanes());
  c->lanes.clear();
  for (int l = 0; l < c->NL; ++l) {
    c->lanes.emplace_back(new Lane());
    Lane& ln = *c->lanes.back();
    if (l == 0) {
      ln.stream = c->stream;
    } else if (c->batch && l % c->batch != 0) {
      ln.stream = c->lanes[(size_t)(l / c->batch * c->batch)]->stream;   // the batch's stream
    } else {
      HIP_CHECK(hipStreamCreateWithFlags(&ln.stream, hipStreamNonBlocking));
      ln.own_stream = true;
    }
    HIP_CHECK(hipEventCreateWithFlags(&ln.done_ev, hipEventDisableTiming));
    for (int b = 0; b < 3; ++b) ln.gh[b].alloc((size_t)N * (quant ? 2 : 8));
    ln.nodes.alloc(L);
    ln.leaf_shift.alloc((size_t)2 * L);
    ln.units.alloc(2 * kMaxBatch + 2);   // first units, total, item count, each item's slab form
    ln.units.zero(c->stream);
    ln.acc_units.alloc(kMaxBatch + 1);
    if (quant) ln.ipool.alloc((size_t)MS * F * B);
    else ln.pool.alloc((size_t)MS * F * B * 2);
    ln.feat_split.alloc((size_t)2 * kMaxBatch * F);
    ln.best.alloc(MS);
    ln.log.alloc(L);
    ln.spl.alloc((size_t)MS * F);
    ln.st.alloc(1);
    ln.st.zero(c->stream);
    ln.plan.alloc(1);
    ln.plan.zero(c->stream);
    ln.reps.alloc(L);
    ln.pos.alloc((size_t)2 * L);
    ln.leaves.alloc(MS);
    ln.leaves.zero(c->stream);
    ln.renew.alloc((size_t)2 * L);
    if (quant && o.quant_train_renew_leaf) ln.leaf16.alloc(N);
    ln.renew.zero(c->stream);
  }
  if (!c->grad_ev) HIP_CHECK(hipEventCreateWithFlags(&c->grad_ev, hipEventDisableTiming));
  c->it.alloc(K);
  c->it.zero(c->stream);
  c->leafval.alloc((size_t)K * L);
  c->leafraw.alloc((size_t)K * L);
  set_part_lds_limit(part_lds(L));
  if (quant) {
    // quant.rs:45-79: per-row tables from block generators of `quant_table_threads` blocks
    std::vector<double> grv(N), hrv(N);
    int threads = std::max(1, o.quant_table_threads);
    int64_t nblock = std::max<int64_t>(1, std::min<int64_t>(threads, (N + 511) / 512));
    int64_t bsize = N;
    if (nblock > 1) { bsize = (N + nblock - 1) / nblock; bsize = (bsize + 31) / 32 * 32; }
    for (int64_t b = 0; b < nblock; ++b) {
      Mt ge((uint32_t)(o.seed + (int)b)), he((uint32_t)(o.seed + (int)b + threads));
      for (int64_t i = b * bsize; i < std::min(N, (b + 1) * bsize); ++i) {
        grv[i] = ge.uniform01();
        hrv[i] = he.uniform01();
      }
    }
    c->grv.alloc(N);
    c->hrv.alloc(N);
    HIP_CHECK(hipMemcpy(c->grv.p, grv.data(), N * 8, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(c->hrv.p, hrv.data(), N * 8, hipMemcpyHostToDevice));
    c->start_eng.reset(new Mt((uint32_t)o.seed));
    const int cap = 1 << 16;
    std::vector<int> s(cap);
    for (int i = 0; i < cap; ++i) s[i] = c->start_eng->uniform_int(0, (int)N);
    c->starts.alloc(cap);
    HIP_CHECK(hipMemcpy(c->starts.p, s.data(), cap * 4, hipMemcpyHostToDevice));
    c->call_ctr.alloc(1);
    c->call_ctr.zero(c->stream);
    c->gq.alloc(N);
    c->hq.alloc(N);
    c->n_calls = 0;
  }
  set_shapes(c, quant);
  {
    const size_t gbytes = (size_t)c->hist_ng * N * c->hist_fb;
    // row bagging: Bagger setup (sample_strategy.rs:186-214); the feature
    // count stands for the reference's feature groups (dense data)
    // balanced bagging needs the objective's positive count (binary.rs:73-99; 0 otherwise)
    int64_t num_pos = 0;
    if (o.objective == 0)
      for (float l : c->label) num_pos += l > 0.0f ? 1 : 0;
    c->bag_balanced = (o.pos_bagging_fraction < 1.0 || o.neg_bagging_fraction < 1.0) && num_pos > 0;
    c->bag_on = (o.bagging_fraction < 1.0 || c->bag_balanced) && o.bagging_freq > 0;
    c->goss_on = o.data_sample_strategy == 1;
    c->bag_active = false;
    c->bag_cnt = N;
    if (c->goss_on) {
      // GOSSStrategy::ResetSampleConfig (sample_strategy.rs:140-160)
      if (o.top_rate + o.other_rate > 1.0)
        throw Error(LGBMHIP_INVALID_PARAMETER, "Check failed: (config_->top_rate + config_->other_rate) <= (1.0f)");
      if (!(o.top_rate > 0.0 && o.other_rate > 0.0))
        throw Error(LGBMHIP_INVALID_PARAMETER, "Check failed: config_->top_rate > 0.0f && config_->other_rate > 0.0f");
      if (o.bagging_freq > 0 && o.bagging_fraction != 1.0) throw Error(LGBMHIP_INVALID_PARAMETER, "Cannot use bagging in GOSS");
      c->bag_subset = o.top_rate + o.other_rate <= 0.5;

      c->bag_on = false;
      // threading.rs block_info_force_size(num_threads, n, 1024)
      const int64_t T = std::max(1, o.sample_threads);
      int64_t nblock = std::min<int64_t>(T, (N + 1023) / 1024), block = N;
      if (nblock > 1) {
        block = (N + nblock - 1) / nblock;
        block = (block + 1023) / 1024 * 1024;
      }
      std::vector<GossBlock> gbs;
      for (int64_t b = 0; b < nblock; ++b) {
        const int64_t start = b * block;
        if (start >= N) break;
        GossBlock gb{};
        gb.start = start;
        gb.cnt = std::min(block, N - start);
        gb.top_k = std::max((int)((double)gb.cnt * o.top_rate), 1);
        gb.other_k = (int)((double)gb.cnt * o.other_rate);
        gb.multiply = (float)(int)(gb.cnt - gb.top_k) / (float)gb.other_k;
        gbs.push_back(gb);
      }
      c->goss_nblock = (int)gbs.size();
      c->goss_blocks.alloc(gbs.size());
      HIP_CHECK(hipMemcpy(c->goss_blocks.p, gbs.data(), gbs.size() * sizeof(GossBlock), hipMemcpyHostToDevice));
      c->goss_abs.alloc(N);
      c->goss_thr.alloc(gbs.size());
      const int64_t nb = (N + 1023) / 1024;
      c->bag_rows.alloc(N);
      c->bag_flags.alloc(N);
      c->bag_counts.alloc(nb);
      c->bag_offsets.alloc(nb);
      c->bag_total.alloc(1);
      c->bag_grp0.alloc(gbytes);
      std::vector<unsigned> lcg(nb);
      for (int64_t b = 0; b < nb; ++b) lcg[b] = (unsigned)o.bagging_seed + (unsigned)b;
      c->bag_lcg.alloc(nb);
      HIP_CHECK(hipMemcpy(c->bag_lcg.p, lcg.data(), nb * 4, hipMemcpyHostToDevice));
    }
    if (c->bag_on) {
      c->bag_cnt = c->bag_balanced ? (int64_t)((double)num_pos * o.pos_bagging_fraction) +
                                         (int64_t)((double)(N - num_pos) * o.neg_bagging_fraction)
                                   : (int64_t)(o.bagging_fraction * (double)N);
      // upstream's subset mode counts feature groups, not features (sample_strategy.rs:198-201)
      const int groups = c->num_groups > 0 ? c->num_groups : F;
      c->bag_subset = ((double)c->bag_cnt / (double)N) / (double)o.bagging_freq <= 0.5 && groups < 100;

      c->need_re_bagging = true;
      c->bag_by_query = o.bagging_by_query != 0;
      if (c->bag_by_query) {
        // sample_strategy.rs:239-276; without query data every bag would be empty
        if (c->qb.size() < 2) throw Error(LGBMHIP_UNSUPPORTED, "bagging_by_query without query boundaries");
        if (c->qb.front() != 0 || c->qb.back() != (int32_t)N)
          throw Error(LGBMHIP_INVALID_DATA, "query boundaries do not cover the data");
        c->bag_qb.alloc(c->qb.size());
        HIP_CHECK(hipMemcpy(c->bag_qb.p, c->qb.data(), c->qb.size() * 4, hipMemcpyHostToDevice));
      }
      const int64_t nb = (N + 1023) / 1024;
      c->bag_rows.alloc(N);
      c->bag_flags.alloc(N);
      c->bag_counts.alloc(nb);
      c->bag_offsets.alloc(nb);
      c->bag_total.alloc(1);
      c->bag_grp0.alloc(gbytes);
      c->bag_m = 0;
      // one generator per 1024-row block, Random(bagging_seed + block)
      std::vector<unsigned> lcg(nb);
      for (int64_t b = 0; b < nb; ++b) lcg[b] = (unsigned)o.bagging_seed + (unsigned)b;
      c->bag_lcg.alloc(nb);
      HIP_CHECK(hipMemcpy(c->bag_lcg.p, lcg.data(), nb * 4, hipMemcpyHostToDevice));
    }
    if (quant && (c->bag_on || c->goss_on) && !c->bag_subset) {
      // a full-data bag: the quantizer runs over every row, and the root's
      // f32-scaled sums (quant.rs:144-158) are summed exactly; the
      // reference's f64 running sums are exact (and so equal them) while
      // rows x quantization bins < 2^29
      if ((double)N * (double)std::max(2, o.num_grad_quant_bins) >= 536870912.0)
        throw Error(LGBMHIP_UNSUPPORTED, "quantized gradients with a full-data bag above 2^29 / num_grad_quant_bins rows");
      for (auto& lp : c->lanes) lp->q_all.alloc(N);
    }
    // buffer 0: the root's rows in row order, bins group-major
    c->grp0.alloc(gbytes);
    launch_group_rows(c, nullptr, N, c->grp0.p, c->grp0.n, c->stream);
  }
  for (auto p : c->h_log) if (p) (void)hipHostFree(p);
  for (auto p : c->h_raw) if (p) (void)hipHostFree(p);
  for (auto p : c->h_st) if (p) (void)hipHostFree(p);
  c->h_log.assign(K, nullptr);
  c->h_raw.assign(K, nullptr);
  c->h_st.assign(K, nullptr);
  for (int k = 0; k < K; ++k) {
    HIP_CHECK(hipHostMalloc((void**)&c->h_log[k], sizeof(SplitRec) * L));
    HIP_CHECK(hipHostMalloc((void**)&c->h_raw[k], sizeof(double) * L));
    HIP_CHECK(hipHostMalloc((void**)&c->h_st[k], sizeof(StepDev)));
  }
  drop_graphs(c);
  c->graphs.assign(K, {});
  c->graphs_cont.assign(K, {});
  c->graphs_batch.assign(2 * std::max(1, c->nbatch), {});
  if (c->batch) {
    c->tab_part.alloc(K);
    c->tab_hist_root.alloc(K);
    c->tab_hist_round.alloc(K);
    c->tab_split_root.alloc(K);
    c->tab_split_round.alloc(K);
    c->tab_split_raw.alloc(K);
    c->tab_slab.alloc(K);
  }
  c->last_rounds.assign(K, 0);
  c->want_rounds.assign(K, 0);
  c->trees.clear();
  c->iter = 0;
  c->pred_dirty = true;
  for (auto& v : c->valid) reset_valid_scores(c, *v);
  HIP_CHECK(hipStreamSynchronize(c->stream));
  c->state_ready = true;
}

static HistArgs hist_args(Ctx* c, Lane& ln, bool quant, int k, int mode) {
  HistArgs a{};
  for (int b = 0; b < 3; ++b) {
    a.grp[b] = b == 0 ? (c->bag_active ? c->bag_grp0.p : c->grp0.p) : ln.grp[b].p;
    a.gh[b] = ln.gh[b].p;
  }
  a.st = ln.st.p;
  a.plan = ln.plan.p;
  a.reps = ln.reps.p;
  a.leaves = ln.leaves.p;
  a.log = ln.log.p;
  a.best = ln.best.p;
  a.feats = c->d_feats.p;
  a.bytes = c->d_bytes.p;
  a.ginfo = c->d_ginfo.p;
  a.F = c->F;
  a.B = c->B;
  a.FB = c->hist_fb;
  a.NG = c->hist_ng;
  a.pu = c->hist_pu;
  a.u4 = c->has_u4 ? 1 : 0;
  a.L = c->L;
  a.N = c->N;
  a.max_chunks = c->hist_chunks;
  a.min_rows_per_chunk = c->hist_min_rows;
  a.slab = ln.slab.p;
  a.units = ln.units.p;
  a.it = c->it.p + k;
  a.p = c->sp;
  a.quant = quant ? 1 : 0;
  a.mode = mode;
  a.root_n = c->bag_active ? c->bag_m : c->N;
  return a;
}

static SplitArgs split_args(Ctx* c, Lane& ln, int k, bool root) {
  SplitArgs a{};
  a.st = ln.st.p;
  a.plan = ln.plan.p;
  a.leaves = ln.leaves.p;
  a.feats = c->d_feats.p;
  a.p = c->sp;
  a.F = c->F;
  a.B = c->B;
  a.L = c->L;
  a.slab = ln.acc.p;          // summed by the reduction after each histogram pass
  a.units = ln.acc_units.p;
  a.FB = c->hist_fb;
  a.NG = c->hist_ng;
  a.pu = c->hist_pu;
  a.pool = ln.pool.p;
  a.ipool = ln.ipool.p;
  a.feat_split = ln.feat_split.p;
  a.splittable = ln.spl.p;
  a.it = c->it.p + k;
  a.quant = c->opts.use_quantized_grad != 0;
  a.root = root ? 1 : 0;
  return a;
}

static PartArgs part_args(Ctx* c, Lane& ln) {
  PartArgs a{};
  a.st = ln.st.p;
  a.plan = ln.plan.p;
  a.reps = ln.reps.p;
  a.pos = ln.pos.p;
  a.leaves = ln.leaves.p;
  a.feats = c->d_feats.p;
  a.best = ln.best.p;
  a.feat_split = ln.feat_split.p;
  for (int b = 0; b < 3; ++b) {
    a.grp[b] = b == 0 ? (c->bag_active ? c->bag_grp0.p : c->grp0.p) : ln.grp[b].p;
    a.gh[b] = ln.gh[b].p;
  }
  a.FB = c->hist_fb;
  a.NG = c->hist_ng;
  a.N = c->N;
  a.gh_bytes = c->opts.use_quantized_grad ? 2 : 8;
  a.F = c->F;
  a.L = c->L;
  return a;
}

// Profiling events: per class, step 0 = the root's histogram and search,
// step r + 1 = growth round r; kinds 0 histogram, 1 search, 2 partition,
// 3 slab reduction.
static int prof_steps(const Ctx* c) { return c->L + 2; }
static hipEvent_t pevent(Ctx* c, int k, int step, int kind, int end) {
  size_t i = ((((size_t)k * prof_steps(c) + step) * 4 + kind) * 2 + end);
  return c->pev[i];
}
static void prof_mark(Ctx* c, int k, int step, int kind, int end) {
  if (c->profile && step < prof_steps(c)) HIP_CHECK(hipEventRecord(pevent(c, k, step, kind, end), lane_of(c, k).stream));
}

// Every item's histogram units of the last histogram pass summed into acc
// (one unit per item and feature group).
static void slab_total(Ctx* c, Lane& ln, bool quant, hipStream_t s) {
  const int per_unit = c->hist_pu * (quant ? 1 : 2);
  const int64_t eblocks = ((int64_t)c->hist_ng * per_unit + 31) / 32;
  const int grid = (int)std::min<int64_t>(eblocks * kMaxBatch, 4LL * c->cus);
  hipLaunchKernelGGL(k_slab_reduce, dim3(grid), dim3(256), 0, s, ln.slab.p, ln.units.p, c->hist_ng, per_unit,
                     ln.acc.p);
}

// The launches that open a tree for class k (gradients already in g/h and
// max |g|, |h| in it[k]): identity partition or quantization with the root
// sums, and the growth state.
static void enqueue_head_pre(Ctx* c, Lane& ln, int k) {
  hipStream_t s = ln.stream;
  const int64_t N = c->N;
  const bool quant = c->opts.use_quantized_grad != 0;
  const float* gk = c->g.p + (size_t)k * N;
  const float* hk = c->h.p + (size_t)k * N;
  // the root: every row, or the bag's rows (positions 0 .. bag_m - 1)
  const int64_t R = c->bag_active ? c->bag_m : N;
  const unsigned* bag = c->bag_active ? c->bag_rows.p : nullptr;
  const int ig = grid_for(R, 256, c->cus * 4);
  // a quantized tree on a full-data bag: the quantizer over every row, the
  // root's positions and sums from the bag (quant.rs:92-158)
  const bool full_bag = quant && bag && !c->bag_subset;
  if (full_bag) {
    const int ag = grid_for(N, 256, c->cus * 4);
    hipLaunchKernelGGL(k_clear_max, dim3(1), dim3(1), 0, s, c->it.p + k);
    hipLaunchKernelGGL(k_absmax_it, dim3(ag), dim3(256), 0, s, gk, hk, (long long)N, c->it.p + k,
                       (const unsigned*)nullptr);
    hipLaunchKernelGGL(k_discretize, dim3(ag), dim3(256), 0, s, (long long)N, gk, hk, c->grv.p, c->hrv.p,
                       c->opts.num_grad_quant_bins, c->opts.stochastic_rounding, 0, c->starts.p, c->call_ctr.p, k,
                       c->it.p + k, (unsigned*)nullptr, (char2*)ln.q_all.p, (int8_t*)nullptr, (int8_t*)nullptr,
                       (const unsigned*)nullptr);
    hipLaunchKernelGGL(k_full_bag_root, dim3(ig), dim3(256), 0, s, (long long)R, bag, (const char2*)ln.q_all.p,
                       c->it.p + k, (char2*)ln.gh[0].p);
  } else if (bag) {
    // max |g|, |h| over the bag (quant.rs:100-108 with rows)
    hipLaunchKernelGGL(k_clear_max, dim3(1), dim3(1), 0, s, c->it.p + k);
    hipLaunchKernelGGL(k_absmax_it, dim3(ig), dim3(256), 0, s, gk, hk, (long long)R, c->it.p + k, bag);
  }
  if (full_bag) {
  } else if (quant) {
    // the k-th discretizer call of the iteration (the counter moves on by
    // the iteration's trees once they are all grown)
    hipLaunchKernelGGL(k_discretize, dim3(ig), dim3(256), 0, s, (long long)R, gk, hk, c->grv.p, c->hrv.p,
                       c->opts.num_grad_quant_bins, c->opts.stochastic_rounding, 0, c->starts.p, c->call_ctr.p, k,
                       c->it.p + k, (unsigned*)nullptr, (char2*)ln.gh[0].p, (int8_t*)nullptr, (int8_t*)nullptr, bag);
  } else {
    hipLaunchKernelGGL(k_tree_init, dim3(ig), dim3(256), 0, s, (long long)R, gk, hk, (unsigned*)nullptr,
                       (float2*)ln.gh[0].p, c->it.p + k, bag);
  }
  hipLaunchKernelGGL(k_state_init, dim3(1), dim3(256), 0, s, ln.st.p, ln.pos.p, ln.leaves.p, ln.spl.p, (long long)R,
                     c->F);
}

// The first launches of a tree for class k: enqueue_head_pre, then the
// root's histogram and search.
static void enqueue_head(Ctx* c, Lane& ln, int k) {
  hipStream_t s = ln.stream;
  const bool quant = c->opts.use_quantized_grad != 0;
  enqueue_head_pre(c, ln, k);
  const int hgrid = c->hist_grid;
  prof_mark(c, k, 0, 0, 0);
  launch_hist(hist_args(c, ln, quant, k, 2), quant, c->hist_tpr, hgrid, c->hist_lds, s);
  prof_mark(c, k, 0, 0, 1);
  slab_total(c, ln, quant, s);
  prof_mark(c, k, 0, 3, 1);
  launch_split(split_args(c, ln, k, true), s);
  prof_mark(c, k, 0, 1, 1);
}

// Growth rounds r0 .. r0 + n - 1 (a round after the tree is done returns at once).
static void enqueue_rounds(Ctx* c, Lane& ln, int k, int r0, int n) {
  hipStream_t s = ln.stream;
  const bool quant = c->opts.use_quantized_grad != 0;
  const HistArgs ha = hist_args(c, ln, quant, k, 3);
  const SplitArgs sa = split_args(c, ln, k, false);
  // rounds from late_round on: few, small leaves of one or two row chunks,
  // whose units the search sums itself, one launch fewer per round
  SplitArgs sa_raw = sa;
  sa_raw.slab = ln.slab.p;
  sa_raw.units = ln.units.p;
  sa_raw.raw = 1;
  const PartArgs pa = part_args(c, ln);
  const int hgrid = c->hist_grid;
  for (int r = r0; r < r0 + n; ++r) {
    const int step = r + 1;
    const bool late = c->late_round > 0 && r >= c->late_round;
    launch_plan(pa, s);
    prof_mark(c, k, step, 2, 0);
    launch_partition(pa, c->part_grid, s);
    prof_mark(c, k, step, 2, 1);
    launch_hist(ha, quant, c->hist_tpr, hgrid, c->hist_lds, s);
    prof_mark(c, k, step, 0, 1);
    if (!late) slab_total(c, ln, quant, s);
    prof_mark(c, k, step, 3, 1);
    launch_split(late ? sa_raw : sa, s);
    prof_mark(c, k, step, 1, 1);
  }
}

// The tree's end: quantized leaf renewal, leaf values, score update. No-ops
// while the tree is still growing.
static WalkArgs walk_args(Ctx* c, Lane& ln, int mode) {
  WalkArgs w{};
  w.st = ln.st.p;
  w.nodes = ln.nodes.p;
  w.feats = c->d_feats.p;
  w.bins = c->bins_rm.p;
  w.S = c->S;
  w.L = c->L;
  w.N = c->N;
  w.shifts = ln.leaf_shift.p;
  w.sums = ln.renew.p;
  w.mode = mode;
  return w;
}

// The tree's end: its nodes, quantized leaf renewal, leaf values, score
// update (the last two walk the tree over every row's bins in row order).
// No-ops while the tree is still growing.
static void enqueue_tail(Ctx* c, Lane& ln, int k, bool score_update, double lr) {
  hipStream_t s = ln.stream;
  const int64_t N = c->N;
  const bool quant = c->opts.use_quantized_grad != 0;
  const bool renew = quant && c->opts.quant_train_renew_leaf;
  const int wgrid = grid_for(N, 256, c->cus * 4);
  hipLaunchKernelGGL(k_tree_nodes, dim3(1), dim3(256), 0, s, ln.st.p, ln.log.p, ln.pos.p, ln.leaves.p, c->it.p + k,
                     c->L, ln.nodes.p, ln.leaf_shift.p);
  if (renew) {
    // leaf sums over the root's rows: every row, or the bag's
    WalkArgs w = walk_args(c, ln, 0);
    w.g = c->g.p + (size_t)k * N;
    w.h = c->h.p + (size_t)k * N;
    if (c->bag_active) {
      w.rows = c->bag_rows.p;
      w.N = c->bag_m;
    } else {
      w.leaf16 = ln.leaf16.p;   // every row's leaf, for the score update below
    }
    launch_walk(w, c->bag_active ? grid_for(c->bag_m, 256, c->cus * 4) : wgrid, s);
  }
  hipLaunchKernelGGL(k_leaf_values, dim3((c->L + 255) / 256), dim3(256), 0, s, ln.st.p, ln.pos.p, ln.leaves.p,
                     c->it.p + k, ln.renew.p, renew ? 1 : 0, c->sp, lr, c->L, c->leafval.p + (size_t)k * c->L,
                     c->leafraw.p + (size_t)k * c->L);
  if (score_update) {
    WalkArgs w = walk_args(c, ln, 1);
    w.val = c->leafval.p + (size_t)k * c->L;
    w.score = c->score.p + (size_t)k * N;
    if (renew && !c->bag_active)
      hipLaunchKernelGGL(k_score_leaf, dim3(grid_for(N, 1024, c->cus * 4)), dim3(256), 0, s, ln.st.p, ln.leaf16.p,
                         w.val, w.score, (long long)N);
    else
      launch_walk(w, wgrid, s);
    // every validation set: the same walk over its bins (tree.rs:635-652)
    for (auto& v : c->valid) {
      WalkArgs wv = w;
      wv.bins = v->bins.p;
      wv.N = v->n;
      wv.score = v->score.p + (size_t)k * v->n;
      launch_walk(wv, grid_for(v->n, 256, c->cus * 4), s);
    }
  }
  hipLaunchKernelGGL(k_tail_done, dim3(1), dim3(64), 0, s, ln.st.p);
}

static hipGraphExec_t capture(hipStream_t s, const std::function<void()>& body) {
  hipGraph_t g;
  HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  body();
  HIP_CHECK(hipStreamEndCapture(s, &g));
  hipGraphExec_t e = nullptr;
  HIP_CHECK(hipGraphInstantiate(&e, g, nullptr, nullptr, 0));
  (void)hipGraphDestroy(g);
  return e;
}

// Rounds of class k's next main launch: at least rounds_main, else what its
// previous tree ran to completion plus two, rounded up to a multiple of 8
// (consecutive trees of a class take similar round counts, so most trees
// finish in one launch; rounds after the tree is done do nothing).
static int main_rounds(const Ctx* c, int k) {
  const int want = c->want_rounds[k] > 0 ? c->want_rounds[k] + 2 : 0;
  return std::max(c->rounds_main, (want + 7) / 8 * 8);
}

// Class k's tree on its lane: the main launch (graph or stream: init, root,
// `rounds` rounds, tail), then a copy of the growth state for the host.
static void launch_main(Ctx* c, int k, int rounds, bool score_update, double lr) {
  Lane& ln = lane_of(c, k);
  // a bag changes the root's size and rows: stream launches
  const bool graph = c->opts.use_graph && !c->profile && !c->bag_on && !c->goss_on;
  if (graph) {
    const Ctx::GraphKey key{rounds, score_update, lr};
    hipGraphExec_t e = nullptr;
    for (auto& g : c->graphs[k])
      if (g.first == key) e = g.second;
    if (!e) {
      e = capture(ln.stream, [&] {
        enqueue_head(c, ln, k);
        enqueue_rounds(c, ln, k, 0, rounds);
        enqueue_tail(c, ln, k, score_update, lr);
      });
      c->graphs[k].emplace_back(key, e);
    }
    HIP_CHECK(hipGraphLaunch(e, ln.stream));
  } else {
    enqueue_head(c, ln, k);
    enqueue_rounds(c, ln, k, 0, rounds);
    enqueue_tail(c, ln, k, score_update, lr);
  }
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(c->h_st[k], ln.st.p, sizeof(StepDev), hipMemcpyDeviceToHost, ln.stream));
}

// rounds_cont more rounds of class k's tree and its tail again.
static void launch_cont(Ctx* c, int k, int launched, bool score_update, double lr) {
  Lane& ln = lane_of(c, k);
  const bool graph = c->opts.use_graph && !c->profile && !c->bag_on && !c->goss_on;
  if (graph) {
    const Ctx::GraphKey key{c->rounds_cont, score_update, lr};
    hipGraphExec_t e = nullptr;
    for (auto& g : c->graphs_cont[k])
      if (g.first == key) e = g.second;
    if (!e) {
      e = capture(ln.stream, [&] {
        enqueue_rounds(c, ln, k, 0, c->rounds_cont);
        enqueue_tail(c, ln, k, score_update, lr);
      });
      c->graphs_cont[k].emplace_back(key, e);
    }
    HIP_CHECK(hipGraphLaunch(e, ln.stream));
  } else {
    enqueue_rounds(c, ln, k, launched, c->rounds_cont);
    enqueue_tail(c, ln, k, score_update, lr);
  }
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(c->h_st[k], ln.st.p, sizeof(StepDev), hipMemcpyDeviceToHost, ln.stream));
}

// ---------------------------------------------------------------- batched trees
// The class trees of an iteration grow in batches: every growth launch
// (plans, partitions, histograms, slab sums, searches) carries all trees of
// a batch, each kernel taking its tree's arguments from a table in device
// memory (grid y, or z for the search, is the tree). Each tree's blocks
// loop over its pending leaves of uneven row counts as in a single tree's
// launches; per-tree grids are the tuned ones divided among the batch's
// trees. Batches run on their own streams, like growth lanes.
static void fill_tabs(Ctx* c) {
  const int K = c->K;
  const bool quant = c->opts.use_quantized_grad != 0;
  c->h_tab_part.resize(K);
  c->h_tab_hist_root.resize(K);
  c->h_tab_hist_round.resize(K);
  c->h_tab_split_root.resize(K);
  c->h_tab_split_round.resize(K);
  c->h_tab_split_raw.resize(K);
  c->h_tab_slab.resize(K);
  for (int k = 0; k < K; ++k) {
    Lane& ln = *c->lanes[k];
    c->h_tab_part[k] = part_args(c, ln);
    c->h_tab_hist_root[k] = hist_args(c, ln, quant, k, 2);
    c->h_tab_hist_round[k] = hist_args(c, ln, quant, k, 3);
    c->h_tab_split_root[k] = split_args(c, ln, k, true);
    c->h_tab_split_round[k] = split_args(c, ln, k, false);
    // rounds from late_round on: the search reads the histogram units
    c->h_tab_split_raw[k] = c->h_tab_split_round[k];
    c->h_tab_split_raw[k].slab = ln.slab.p;
    c->h_tab_split_raw[k].units = ln.units.p;
    c->h_tab_split_raw[k].raw = 1;
    c->h_tab_slab[k] = SlabTab{ln.slab.p, ln.units.p, ln.acc.p};
  }
  hipStream_t s = c->stream;
  HIP_CHECK(hipMemcpyAsync(c->tab_part.p, c->h_tab_part.data(), sizeof(PartArgs) * K, hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(c->tab_hist_root.p, c->h_tab_hist_root.data(), sizeof(HistArgs) * K,
                           hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(c->tab_hist_round.p, c->h_tab_hist_round.data(), sizeof(HistArgs) * K,
                           hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(c->tab_split_root.p, c->h_tab_split_root.data(), sizeof(SplitArgs) * K,
                           hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(c->tab_split_round.p, c->h_tab_split_round.data(), sizeof(SplitArgs) * K,
                           hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(c->tab_split_raw.p, c->h_tab_split_raw.data(), sizeof(SplitArgs) * K,
                           hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(c->tab_slab.p, c->h_tab_slab.data(), sizeof(SlabTab) * K, hipMemcpyHostToDevice, s));
}

// Batch g: its first class, its tree count and its stream.
struct BatchRef {
  int k0, T;
  hipStream_t s;
};
static BatchRef batch_ref(Ctx* c, int g) {
  const int k0 = g * c->batch;
  return BatchRef{k0, std::min(c->batch, c->K - k0), c->lanes[(size_t)k0]->stream};
}

// Blocks per tree of a batched launch: the single tree's grid shared among
// the batch's T trees (at least `floor`).
static int batch_grid(int grid, int T, int floor) { return std::max(floor, grid / T); }

static void slab_total_batch(Ctx* c, const BatchRef& b, bool quant) {
  const int per_unit = c->hist_pu * (quant ? 1 : 2);
  const int64_t eblocks = ((int64_t)c->hist_ng * per_unit + 31) / 32;
  const int grid = (int)std::min<int64_t>(eblocks * kMaxBatch, 4LL * c->cus);
  hipLaunchKernelGGL(k_slab_reduce_tab, dim3(batch_grid(grid, b.T, 32), b.T), dim3(256), 0, b.s,
                     c->tab_slab.p + b.k0, c->hist_ng, per_unit);
}

// Every tree's opening launches, then the batch's roots' histograms and searches.
static void enqueue_head_batch(Ctx* c, const BatchRef& b) {
  const bool quant = c->opts.use_quantized_grad != 0;
  for (int k = b.k0; k < b.k0 + b.T; ++k) enqueue_head_pre(c, *c->lanes[k], k);
  prof_mark(c, b.k0, 0, 0, 0);
  launch_hist_tab(c->tab_hist_root.p + b.k0, b.T, quant, c->has_u4, c->hist_tpr, batch_grid(c->hist_grid, b.T, 32),
                  c->hist_lds, b.s);
  prof_mark(c, b.k0, 0, 0, 1);
  slab_total_batch(c, b, quant);
  prof_mark(c, b.k0, 0, 3, 1);
  launch_split_tab(c->tab_split_root.p + b.k0, b.T, c->F, true, b.s);
  prof_mark(c, b.k0, 0, 1, 1);
}

static void enqueue_rounds_batch(Ctx* c, const BatchRef& b, int r0, int n) {
  const bool quant = c->opts.use_quantized_grad != 0;
  const PartArgs& pa = c->h_tab_part[(size_t)b.k0];
  const int pgrid = batch_grid(c->part_grid, b.T, 32), hgrid = batch_grid(c->hist_grid, b.T, 32);
  for (int r = r0; r < r0 + n; ++r) {
    const int step = r + 1;
    launch_plan_tab(c->tab_part.p + b.k0, b.T, pa, b.s);
    prof_mark(c, b.k0, step, 2, 0);
    launch_partition_tab(c->tab_part.p + b.k0, b.T, pa, pgrid, b.s);
    prof_mark(c, b.k0, step, 2, 1);
    launch_hist_tab(c->tab_hist_round.p + b.k0, b.T, quant, c->has_u4, c->hist_tpr, hgrid, c->hist_lds, b.s);
    prof_mark(c, b.k0, step, 0, 1);
    const bool late = c->late_round > 0 && r >= c->late_round;
    if (!late) slab_total_batch(c, b, quant);
    prof_mark(c, b.k0, step, 3, 1);
    launch_split_tab((late ? c->tab_split_raw.p : c->tab_split_round.p) + b.k0, b.T, c->F, false, b.s);
    prof_mark(c, b.k0, step, 1, 1);
  }
}

// One launch (a captured graph when graphs are on) of batch g: its trees'
// opening launches when `head`, `rounds` batched rounds, every tree's tail;
// then copies of the batch's growth states for the host.
static void launch_batch(Ctx* c, int g, bool head, int rounds, int r0, bool score_update, double lr) {
  const BatchRef b = batch_ref(c, g);
  auto body = [&](int rr0) {
    if (head) enqueue_head_batch(c, b);
    enqueue_rounds_batch(c, b, rr0, rounds);
    for (int k = b.k0; k < b.k0 + b.T; ++k) enqueue_tail(c, *c->lanes[k], k, score_update, lr);
  };
  const bool graph = c->opts.use_graph && !c->profile && !c->bag_on && !c->goss_on;
  if (graph) {
    const Ctx::GraphKey key{rounds, score_update, lr};
    auto& gs = c->graphs_batch[(size_t)(2 * g + (head ? 0 : 1))];
    hipGraphExec_t e = nullptr;
    for (auto& x : gs)
      if (x.first == key) e = x.second;
    if (!e) {
      e = capture(b.s, [&] { body(0); });
      gs.emplace_back(key, e);
    }
    HIP_CHECK(hipGraphLaunch(e, b.s));
  } else {
    body(r0);
  }
  HIP_CHECK(hipGetLastError());
  for (int k = b.k0; k < b.k0 + b.T; ++k)
    HIP_CHECK(hipMemcpyAsync(c->h_st[k], c->lanes[k]->st.p, sizeof(StepDev), hipMemcpyDeviceToHost, b.s));
}

// The class trees of an iteration in batches: each batch's main launch of as
// many rounds as its classes' previous trees needed, then continuation
// launches while any of its trees still grows (a grown tree's launches
// return at once and its tail is applied once). The host polls the batches
// in turn; the context's stream ends ordered after every batch.
static void grow_trees_batch(Ctx* c, bool score_update, double lr) {
  const int K = c->K, NB = c->nbatch;
  fill_tabs(c);
  HIP_CHECK(hipEventRecord(c->grad_ev, c->stream));
  std::vector<int> launched(NB, 0);
  std::vector<bool> active(NB, true);
  for (int g = 0; g < NB; ++g) {
    const BatchRef b = batch_ref(c, g);
    if (b.s != c->stream) HIP_CHECK(hipStreamWaitEvent(b.s, c->grad_ev, 0));
    int rounds = 0;
    for (int k = b.k0; k < b.k0 + b.T; ++k) rounds = std::max(rounds, main_rounds(c, k));
    launch_batch(c, g, true, rounds, 0, score_update, lr);
    launched[g] = rounds;
  }
  int left = NB;
  while (left > 0) {
    for (int g = 0; g < NB; ++g) {
      if (!active[g]) continue;
      const BatchRef b = batch_ref(c, g);
      HIP_CHECK(hipStreamSynchronize(b.s));
      bool all = true;
      for (int k = b.k0; k < b.k0 + b.T; ++k) all = all && c->h_st[k]->done != 0;
      if (!all) {
        // every round replays or splits at least one leaf of each growing tree
        if (launched[g] > c->L + 1) throw Error(LGBMHIP_INTERNAL, "tree growth did not finish");
        launch_batch(c, g, false, c->rounds_cont, launched[g], score_update, lr);
        launched[g] += c->rounds_cont;
        continue;
      }
      for (int k = b.k0; k < b.k0 + b.T; ++k) {
        Lane& ln = *c->lanes[k];
        c->last_rounds[k] = launched[g];
        c->want_rounds[k] = c->h_st[k]->rounds;
        HIP_CHECK(hipMemcpyAsync(c->h_log[k], ln.log.p, sizeof(SplitRec) * c->L, hipMemcpyDeviceToHost, b.s));
        HIP_CHECK(hipMemcpyAsync(c->h_raw[k], c->leafraw.p + (size_t)k * c->L, sizeof(double) * c->L,
                                 hipMemcpyDeviceToHost, b.s));
      }
      active[g] = false;
      left--;
    }
  }
  for (int g = 1; g < NB; ++g) {
    const BatchRef b = batch_ref(c, g);
    Lane& ln = *c->lanes[(size_t)b.k0];
    HIP_CHECK(hipEventRecord(ln.done_ev, b.s));
    HIP_CHECK(hipStreamWaitEvent(c->stream, ln.done_ev, 0));
  }
  (void)K;
}

// The trees of classes 0 .. ntrees-1 (gradients already enqueued on the
// context's stream), each to completion: lane l grows classes l, l + NL,
// ... in order while the other lanes grow theirs; the host tops up a tree
// with continuation rounds while its device state says it still grows.
// Leaves each tree's log, raw leaf outputs and state in the pinned copies,
// and the context's stream ordered after every lane's last tree.
static void grow_trees(Ctx* c, int ntrees, bool score_update, double lr) {
  if (c->batch && ntrees == c->K) {
    grow_trees_batch(c, score_update, lr);
    return;
  }
  const int NL = std::min(ntrees, c->NL);
  HIP_CHECK(hipEventRecord(c->grad_ev, c->stream));
  for (int l = 1; l < NL; ++l) HIP_CHECK(hipStreamWaitEvent(c->lanes[l]->stream, c->grad_ev, 0));
  std::vector<int> cur(NL, -1), launched(NL, 0);
  int active = 0;
  auto start = [&](int l, int k) {
    cur[l] = k;
    launched[l] = main_rounds(c, k);
    launch_main(c, k, launched[l], score_update, lr);
    active++;
  };
  for (int l = 0; l < NL; ++l) start(l, l);
  while (active > 0) {
    for (int l = 0; l < NL; ++l) {
      const int k = cur[l];
      if (k < 0) continue;
      Lane& ln = *c->lanes[l];
      HIP_CHECK(hipStreamSynchronize(ln.stream));
      if (!c->h_st[k]->done) {
        // every round replays or splits at least one leaf: L + 1 rounds finish any tree
        if (launched[l] > c->L + 1) throw Error(LGBMHIP_INTERNAL, "tree growth did not finish");
        launch_cont(c, k, launched[l], score_update, lr);
        launched[l] += c->rounds_cont;
        continue;
      }
      c->last_rounds[k] = launched[l];
      c->want_rounds[k] = c->h_st[k]->rounds;
      HIP_CHECK(hipMemcpyAsync(c->h_log[k], ln.log.p, sizeof(SplitRec) * c->L, hipMemcpyDeviceToHost, ln.stream));
      HIP_CHECK(hipMemcpyAsync(c->h_raw[k], c->leafraw.p + (size_t)k * c->L, sizeof(double) * c->L,
                               hipMemcpyDeviceToHost, ln.stream));
      cur[l] = -1;
      active--;
      if (k + NL < ntrees) start(l, k + NL);
    }
  }
  for (int l = 1; l < NL; ++l) {
    HIP_CHECK(hipEventRecord(c->lanes[l]->done_ev, c->lanes[l]->stream));
    HIP_CHECK(hipStreamWaitEvent(c->stream, c->lanes[l]->done_ev, 0));
  }
}

static HostTree build_tree(Ctx* c, int k) {
  HostTree t;
  const StepDev& st = *c->h_st[k];
  if (st.num_splits == 0) {
    t.leaf_value[0] = st.root_output;
    t.leaf_count[0] = (int)c->N;
    return t;
  }
  t.leaf_value[0] = st.root_output;
  for (int i = 0; i < st.num_splits; ++i) {
    const SplitRec& r = c->h_log[k][i];
    const int f = r.feature;
    const auto& m = c->metas[f];
    const float gain = (float)(r.gain + c->sp.min_gain_to_split);
    if (m.bin_type == 0) {
      t.split(r.leaf, f, m.real_index, (uint32_t)r.threshold, c->upper[f][r.threshold], r.left_output,
              r.right_output, r.left_count, r.right_count, r.lsh, r.rsh, gain, m.missing_type, r.default_left != 0);
    } else {
      std::vector<uint32_t> bits_in, bits;
      int maxw = -1, maxc = -1;
      for (int b = 0; b < kMaxBins; ++b)
        if ((r.cat_bits[b >> 5] >> (b & 31)) & 1u) {
          maxw = std::max(maxw, b / 32);
          maxc = std::max(maxc, (int)c->upper[f][b] / 32);
        }
      bits_in.assign(maxw + 1, 0);
      bits.assign(maxc + 1, 0);
      for (int b = 0; b < kMaxBins; ++b)
        if ((r.cat_bits[b >> 5] >> (b & 31)) & 1u) {
          bits_in[b / 32] |= 1u << (b % 32);
          int cat = (int)c->upper[f][b];
          bits[cat / 32] |= 1u << (cat % 32);
        }
      t.split(r.leaf, f, m.real_index, (uint32_t)t.num_cat, (double)t.num_cat, r.left_output, r.right_output,
              r.left_count, r.right_count, r.lsh, r.rsh, gain, m.missing_type, false);
      t.decision_type.back() |= 1;
      t.num_cat++;
      t.cat_boundaries.push_back(t.cat_boundaries.back() + (int)bits.size());
      t.cat_threshold.insert(t.cat_threshold.end(), bits.begin(), bits.end());
      t.cat_boundaries_inner.push_back(t.cat_boundaries_inner.back() + (int)bits_in.size());
      t.cat_threshold_inner.insert(t.cat_threshold_inner.end(), bits_in.begin(), bits_in.end());
    }
  }
  if (c->opts.use_quantized_grad && c->opts.quant_train_renew_leaf)
    for (int l = 0; l < t.num_leaves; ++l) t.leaf_value[l] = c->h_raw[k][l];
  return t;
}

static double boost_from_score(Ctx* c, int k) {
  const auto& lab = c->label;
  const int64_t n = (int64_t)lab.size();
  if (c->opts.objective == 0) {
    double suml = 0.0;
    for (int64_t i = 0; i < n; ++i) suml += (double)(int)(lab[i] > 0.0f);
    double pavg = suml / (double)n;
    const double eps = (double)1e-15f;
    pavg = std::min(pavg, 1.0 - eps);
    pavg = std::max(pavg, eps);
    return std::log(pavg / ((double)1.0f - pavg)) / c->opts.sigmoid;
  }
  if (c->opts.objective == 1) {
    double cnt = 0.0;
    for (int64_t i = 0; i < n; ++i) cnt += (int)lab[i] == k ? 1.0 : 0.0;
    return std::log(std::max((double)1e-15f, cnt / (double)n));
  }
  return 0.0;
}

__global__ void k_add_const(double* s, long long n, double v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) s[i] += v;
}


static void enqueue_gradients(Ctx* c) {
  hipStream_t s = c->stream;
  const int64_t N = c->N;
  hipLaunchKernelGGL(k_iter_reset, dim3(1), dim3(64), 0, s, c->it.p, c->K);
  const int grid = grid_for(N, 256, c->cus * 4);
  const float* w = c->has_weight ? c->d_weight.p : nullptr;
  if (c->opts.objective == 0) {
    hipLaunchKernelGGL(k_binary_grad, dim3(grid), dim3(256), 0, s, c->score.p, c->d_label.p, w, (long long)N,
                       c->opts.sigmoid, 1.0, 1.0, c->g.p, c->h.p, c->it.p);
  } else if (c->opts.objective == 2) {
    RankArgs a{};
    a.score = c->score.p;
    a.label = c->d_label.p;
    a.qb = c->d_qb.p;
    a.Q = (int)c->qb.size() - 1;
    a.inv_max_dcg = c->r_imd.p;
    a.table = c->r_table.p;
    a.min_in = c->r_min_in;
    a.max_in = c->r_max_in;
    a.idx_factor = c->r_factor;
    a.sigmoid = c->opts.sigmoid;
    a.norm = c->opts.lambdarank_norm;
    a.trunc = c->opts.lambdarank_truncation_level;
    a.label_gain = c->r_gain.p;
    a.discount = c->r_disc.p;
    a.rank = c->r_rank.p;
    a.order = c->r_order.p;
    a.lam = c->r_lam.p;
    a.hes = c->r_hes.p;
    a.by_rank_s = c->r_bs.p;
    a.by_rank_l = c->r_bl.p;
    a.g = c->g.p;
    a.h = c->h.p;
    if (a.Q > 0) hipLaunchKernelGGL(k_lambdarank, dim3(a.Q), dim3(kRankThreads), 0, s, a);
    hipLaunchKernelGGL(k_absmax_it, dim3(grid), dim3(256), 0, s, c->g.p, c->h.p, (long long)N, c->it.p,
                       (const unsigned*)nullptr);
  } else {
    hipLaunchKernelGGL(k_multiclass_grad, dim3(grid), dim3(256), 0, s, c->score.p, c->d_label.p, w, (long long)N,
                       c->K, c->g.p, c->h.p, c->it.p);
  }
}

// Algorithmic bytes of one histogram build over m rows (row-index read
// skipped at the root, whose rows are contiguous): m * (F + idx + gh) + F*B*out.
// Algorithmic bytes of histogram builds over `rows` rows in `items`
// segments: each row's bins and gradients once (ordered buffers: no index
// read), each segment's histogram written once.
static double hist_bytes(Ctx* c, double rows, double items) {
  const bool quant = c->opts.use_quantized_grad != 0;
  return rows * ((double)c->F + (quant ? 2.0 : 8.0)) + items * (double)c->F * c->B * (quant ? 8.0 : 16.0);
}

// Per-kind device time of the tree just grown (event to event on the
// context's stream, so each figure includes that launch's dispatch gap).
// Every histogram launch counts (one per round plus the root's, as a kernel
// trace counts them).
static void account_profile(Ctx* c, int k) {
  const StepDev& st = *c->h_st[k];
  auto el = [&](hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, a, b));
    return (double)ms;
  };
  c->prof[0] += el(pevent(c, k, 0, 0, 0), pevent(c, k, 0, 0, 1));
  c->prof[1] += 1;
  c->prof[11] += el(pevent(c, k, 0, 0, 1), pevent(c, k, 0, 3, 1));
  c->prof[4] += el(pevent(c, k, 0, 3, 1), pevent(c, k, 0, 1, 1));
  const int steps = std::min(c->last_rounds[k] + 1, prof_steps(c));
  for (int step = 1; step < steps; ++step) {
    c->prof[0] += el(pevent(c, k, step, 2, 1), pevent(c, k, step, 0, 1));
    c->prof[1] += 1;
    c->prof[11] += el(pevent(c, k, step, 0, 1), pevent(c, k, step, 3, 1));
    c->prof[4] += el(pevent(c, k, step, 3, 1), pevent(c, k, step, 1, 1));
    c->prof[5] += el(pevent(c, k, step, 2, 0), pevent(c, k, step, 2, 1));
  }
  c->prof[2] += (double)st.hist_rows;
  c->prof[3] += hist_bytes(c, (double)st.hist_rows, (double)st.hist_items);
  c->prof[8] += (double)st.part_rows;
  c->prof[9] += (double)st.computed;
  c->prof[10] += (double)st.num_splits;
}

// The same for a batch of T trees from class k0: one set of launches (events
// on class k0's slots) for every tree's rows and bytes.
static void account_profile_batch(Ctx* c, int k0, int T) {
  account_profile(c, k0);
  for (int k = k0 + 1; k < k0 + T; ++k) {
    const StepDev& st = *c->h_st[k];
    c->prof[2] += (double)st.hist_rows;
    c->prof[3] += hist_bytes(c, (double)st.hist_rows, (double)st.hist_items);
    c->prof[8] += (double)st.part_rows;
    c->prof[9] += (double)st.computed;
    c->prof[10] += (double)st.num_splits;
  }
}

static void compact_bag(Ctx* c);

// SampleStrategy::bagging (sample_strategy.rs:278-318) on the device: every
// 1024-row block's generator draws its rows (the generators run on from draw
// to draw), the in-bag rows are compacted in ascending order, and buffer 0's
// bins are gathered for them.
static void draw_bag(Ctx* c) {
  const int64_t N = c->N, nb = (N + 1023) / 1024;
  hipStream_t s = c->stream;
  c->need_re_bagging = false;
  if (c->bag_by_query) {
    // whole queries: row flags from the query draws, then the row blocks' counts
    const int64_t nq = (int64_t)c->qb.size() - 1, qbk = (nq + 1023) / 1024;
    hipLaunchKernelGGL(k_bag_draw_queries, dim3((unsigned)((qbk + 3) / 4)), dim3(256), 0, s, (long long)nq,
                       c->bag_lcg.p, c->opts.bagging_fraction, c->bag_qb.p, c->bag_flags.p);
    hipLaunchKernelGGL(k_bag_counts, dim3((unsigned)((nb + 3) / 4)), dim3(256), 0, s, c->bag_flags.p, (long long)N,
                       c->bag_counts.p);
    HIP_CHECK(hipGetLastError());
    compact_bag(c);
    return;
  }
  hipLaunchKernelGGL(k_bag_draw, dim3((unsigned)((nb + 3) / 4)), dim3(256), 0, s, (long long)N,
                     c->bag_lcg.p, c->opts.bagging_fraction, c->bag_balanced ? c->d_label.p : (const float*)nullptr,
                     c->opts.pos_bagging_fraction, c->opts.neg_bagging_fraction, c->bag_flags.p, c->bag_counts.p);
  compact_bag(c);
}

// The in-bag rows (flags set) ascending into bag_rows, their count to the
// host, and the bag's buffer 0 bins.
static void compact_bag(Ctx* c) {
  const int64_t N = c->N, nb = (N + 1023) / 1024;
  hipStream_t s = c->stream;
  hipLaunchKernelGGL(k_bag_offsets, dim3(1), dim3(1024), 0, s, c->bag_counts.p, (long long)nb, c->bag_offsets.p,
                     c->bag_total.p);
  hipLaunchKernelGGL(k_bag_scatter, dim3((unsigned)nb), dim3(1024), 0, s, (long long)N, c->bag_flags.p,
                     c->bag_offsets.p, c->bag_rows.p);
  HIP_CHECK(hipGetLastError());
  long long m = 0;
  HIP_CHECK(hipMemcpyAsync(&m, c->bag_total.p, 8, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  if (m <= 0) throw Error(LGBMHIP_UNSUPPORTED, "an empty bag");
  c->bag_m = m;
  c->bag_cnt = m;
  c->bag_active = true;
  launch_group_rows(c, c->bag_rows.p, m, c->bag_grp0.p, c->bag_grp0.n, s);
}

// GOSSStrategy::Bagging (sample_strategy.rs:319-350, 395-442) after the
// gradients: from iteration 1 / learning_rate on, every iteration's bag.
static void goss_bag(Ctx* c) {
  const int64_t N = c->N, nb = (N + 1023) / 1024;
  hipStream_t s = c->stream;
  c->bag_active = false;
  c->bag_cnt = N;
  if ((int64_t)c->iter < (int64_t)((double)1.0f / c->opts.learning_rate)) return;
  hipLaunchKernelGGL(k_goss_abs, dim3(grid_for(N, 256, c->cus * 4)), dim3(256), 0, s, c->g.p, c->h.p, (long long)N,
                     c->K, c->goss_abs.p);
  hipLaunchKernelGGL(k_goss_threshold, dim3(c->goss_nblock), dim3(1024), 0, s, c->goss_abs.p, c->goss_blocks.p,
                     c->goss_thr.p);
  hipLaunchKernelGGL(k_goss_chain, dim3(c->goss_nblock), dim3(64), 0, s, c->goss_abs.p, c->goss_blocks.p,
                     c->goss_thr.p, c->bag_lcg.p, c->g.p, c->h.p, (long long)N, c->K, c->bag_flags.p);
  hipLaunchKernelGGL(k_bag_counts, dim3((unsigned)((nb + 3) / 4)), dim3(256), 0, s, c->bag_flags.p, (long long)N,
                     c->bag_counts.p);
  HIP_CHECK(hipGetLastError());
  compact_bag(c);
}

// boosting.rs:957-1139 for device=gpu.
static int train_iteration(Ctx* c) {
  if (!c->state_ready) setup_training(c);
  set_device(c);
  const int K = c->K;
  const int64_t N = c->N;
  std::vector<double> init(K, 0.0);
  if (c->trees.empty() && c->opts.boost_from_average) {
    for (int k = 0; k < K; ++k) {
      double v = boost_from_score(c, k);
      if (std::fabs(v) > (double)1e-15f) {
        hipLaunchKernelGGL(k_add_const, dim3(grid_for(N, 256, 2048)), dim3(256), 0, c->stream,
                           c->score.p + (size_t)k * N, (long long)N, v);
        for (auto& vd : c->valid)
          hipLaunchKernelGGL(k_add_const, dim3(grid_for(vd->n, 256, 2048)), dim3(256), 0, c->stream,
                             vd->score.p + (size_t)k * vd->n, (long long)vd->n, v);
        init[k] = v;
      }
    }
  }
  if (!c->ev0) {
    HIP_CHECK(hipEventCreate(&c->ev0));
    HIP_CHECK(hipEventCreate(&c->ev1));
  }
  if (c->opts.use_quantized_grad && c->n_calls + K > (int)c->starts.n)
    throw Error(LGBMHIP_UNSUPPORTED, "more quantized trees than the start-offset table holds (65536)");
  if (c->bag_on && ((c->bag_cnt < N && c->iter % c->opts.bagging_freq == 0) || c->need_re_bagging)) draw_bag(c);
  HIP_CHECK(hipEventRecord(c->ev0, c->stream));
  enqueue_gradients(c);
  if (c->goss_on) goss_bag(c);
  grow_trees(c, K, true, c->opts.learning_rate);
  if (c->opts.use_quantized_grad) {
    hipLaunchKernelGGL(k_bump, dim3(1), dim3(1), 0, c->stream, c->call_ctr.p, K);
    c->n_calls += K;
  }
  HIP_CHECK(hipEventRecord(c->ev1, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
  c->last_ms = ms;
  bool cont = false;
  std::vector<HostTree> made;
  if (c->profile) {
    c->prof[6] += ms;
    c->prof[7] += 1;
  }
  for (int k = 0; k < K; ++k) {
    HostTree t = build_tree(c, k);
    if (c->profile) {
      if (!c->batch) account_profile(c, k);
      else if (k % c->batch == 0) account_profile_batch(c, k, std::min(c->batch, K - k));
    }
    if (t.num_leaves > 1) {
      cont = true;
      t.shrink(c->opts.learning_rate);
      if (std::fabs(init[k]) > (double)1e-15f) t.add_bias(init[k]);
    } else if ((int)c->trees.size() < K) {
      if (!c->opts.boost_from_average) {
        // boosting.rs:1118-1126: a first tree without a split starts the
        // scores from the objective's score
        init[k] = boost_from_score(c, k);
        hipLaunchKernelGGL(k_add_const, dim3(grid_for(N, 256, 2048)), dim3(256), 0, c->stream,
                           c->score.p + (size_t)k * N, (long long)N, init[k]);
        for (auto& vd : c->valid)
          hipLaunchKernelGGL(k_add_const, dim3(grid_for(vd->n, 256, 2048)), dim3(256), 0, c->stream,
                             vd->score.p + (size_t)k * vd->n, (long long)vd->n, init[k]);
        HIP_CHECK(hipStreamSynchronize(c->stream));
      }
      t.as_constant(init[k], (int)N);
    } else {
      t.as_constant(0.0, (int)N);
    }
    made.push_back(t);
  }
  for (auto& t : made) c->trees.push_back(t);
  c->pred_dirty = true;
  if (!cont) {
    if ((int)c->trees.size() > K) c->trees.resize(c->trees.size() - K);
    return 1;
  }
  c->iter++;
  return 0;
}

// ---------------------------------------------------------------- autotuner
// Launch shapes of the tree-growth kernels picked by timing, once per
// workload shape: rows, used features, bins, form, classes, leaves and the
// device's CUs. The reference's only shape rule is its row-block size from
// the thread count (multi_val_bin.rs:337-351 block_info / min_block_size);
// here the knobs are the histogram LDS budget (features per block and the
// group-major row layout), the histogram row chunks and their minimum rows,
// and the partition grid. Trees do not depend on any of them, so candidates
// are timed on the workload itself: each runs boosting iterations from the
// starting scores (two warm ones, which capture its launches, then `trees`
// timed ones) and the training state is reset after it; coordinate descent
// over the knobs in that order.
struct TuneKey {
  int64_t N;
  int F, B, quant, K, L, cus;
  bool operator<(const TuneKey& o) const {
    return std::tie(N, F, B, quant, K, L, cus) < std::tie(o.N, o.F, o.B, o.quant, o.K, o.L, o.cus);
  }
};
struct TuneChoice {
  int part_grid, hist_chunks, hist_min_rows, hist_lds;
};
static std::mutex g_tune_mu;
static std::map<TuneKey, TuneChoice> g_tune_cache;   // per process: one search per workload shape

static void apply_tuning(Ctx* c, const TuneChoice& t) {
  c->tune_part_grid = t.part_grid;
  c->tune_hist_chunks = t.hist_chunks;
  c->tune_hist_min_rows = t.hist_min_rows;
  c->tune_hist_lds = t.hist_lds;
}

// Back to the state before the first iteration: scores (training and
// validation) from zero, no trees, the quantizer's call counter at zero.
static void reset_training(Ctx* c) {
  c->score.zero(c->stream);
  for (auto& v : c->valid) reset_valid_scores(c, *v);
  c->trees.clear();
  c->iter = 0;
  c->pred_dirty = true;
  if (c->opts.use_quantized_grad) {
    c->call_ctr.zero(c->stream);
    c->n_calls = 0;
  }
  std::fill(c->want_rounds.begin(), c->want_rounds.end(), 0);
  HIP_CHECK(hipStreamSynchronize(c->stream));
}

// Time (ms) of `iters` boosting iterations at the current shapes, after
// `warm` iterations that capture their launches, from the starting scores.
static double time_iterations(Ctx* c, int warm, int iters) {
  const bool quant = c->opts.use_quantized_grad != 0;
  set_shapes(c, quant);
  launch_group_rows(c, nullptr, c->N, c->grp0.p, c->grp0.n, c->stream);
  reset_training(c);
  for (int i = 0; i < warm; ++i) train_iteration(c);
  const auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < iters; ++i) train_iteration(c);
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  reset_training(c);
  return ms;
}

// Returns {part_grid, hist_chunks, hist_min_rows, hist_lds_budget, candidates timed}.
static std::array<int, 5> autotune(Ctx* c, int trees) {
  if (!c->state_ready) setup_training(c);
  if (!c->trees.empty()) throw Error(LGBMHIP_INVALID_PARAMETER, "autotune before the first iteration only");
  if (c->bag_on || c->goss_on) throw Error(LGBMHIP_UNSUPPORTED, "autotune with bagging or GOSS");
  set_device(c);
  const bool quant = c->opts.use_quantized_grad != 0;
  const TuneKey key{c->N, c->F, c->B, quant ? 1 : 0, c->K, c->L, c->cus};
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      apply_tuning(c, it->second);
      set_shapes(c, quant);
      launch_group_rows(c, nullptr, c->N, c->grp0.p, c->grp0.n, c->stream);
      HIP_CHECK(hipStreamSynchronize(c->stream));
      return {c->part_grid, c->hist_chunks, c->hist_min_rows, c->tune_hist_lds, 0};
    }
  }
  const int iters = std::max(1, trees);
  const int cus = c->cus;
  TuneChoice best{0, 0, 0, 0};
  apply_tuning(c, best);
  double best_ms = time_iterations(c, 2, iters);
  int timed = 1;
  auto consider = [&](TuneChoice t) {
    apply_tuning(c, t);
    double ms;
    try {
      ms = time_iterations(c, 2, iters);
    } catch (const Error& e) {
      if (e.code != LGBMHIP_UNSUPPORTED) throw;   // a shape this device cannot run
      return;
    }
    ++timed;
    if (ms < best_ms * 0.98) {   // a change must win by more than run-to-run noise
      best_ms = ms;
      best = t;
    }
  };
  // 1. histogram LDS budget (features per block, the row layout): only budgets that change the groups
  {
    std::vector<std::pair<int, int>> seen;
    plan_hist(c, quant);
    seen.emplace_back(c->hist_fb, c->hist_ng);
    for (int kib : {8, 16, 48, 64, 96}) {
      c->tune_hist_lds = kib * 1024;
      plan_hist(c, quant);
      if (std::find(seen.begin(), seen.end(), std::make_pair(c->hist_fb, c->hist_ng)) != seen.end()) continue;
      seen.emplace_back(c->hist_fb, c->hist_ng);
      TuneChoice t = best;
      t.hist_lds = kib * 1024;
      consider(t);
    }
  }
  // 2. histogram row chunks per pass, 3. their minimum rows, 4. partition grid
  for (int m : {1, 2, 4}) {
    TuneChoice t = best;
    t.hist_chunks = m * cus;
    consider(t);
  }
  for (int r : {1024, 2048, 8192}) {
    TuneChoice t = best;
    t.hist_min_rows = r;
    consider(t);
  }
  for (int m : {1, 2, 8}) {
    TuneChoice t = best;
    t.part_grid = m * cus;
    consider(t);
  }
  apply_tuning(c, best);
  set_shapes(c, quant);
  launch_group_rows(c, nullptr, c->N, c->grp0.p, c->grp0.n, c->stream);
  HIP_CHECK(hipStreamSynchronize(c->stream));
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = best;
  }
  return {c->part_grid, c->hist_chunks, c->hist_min_rows, c->tune_hist_lds, timed};
}

// One tree from caller gradients (SerialTreeLearner::train), no score update.
static void train_tree(Ctx* c, const float* g, const float* h) {
  if (!c->state_ready) setup_training(c);
  if (c->bag_on || c->goss_on) throw Error(LGBMHIP_UNSUPPORTED, "one tree from caller gradients with bagging on");
  set_device(c);
  const int64_t N = c->N;
  HIP_CHECK(hipMemcpyAsync(c->g.p, g, N * 4, hipMemcpyHostToDevice, c->stream));
  HIP_CHECK(hipMemcpyAsync(c->h.p, h, N * 4, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_iter_reset, dim3(1), dim3(64), 0, c->stream, c->it.p, c->K);
  hipLaunchKernelGGL(k_absmax_it, dim3(grid_for(N, 256, c->cus * 4)), dim3(256), 0, c->stream, c->g.p, c->h.p,
                     (long long)N, c->it.p, (const unsigned*)nullptr);
  if (c->opts.use_quantized_grad && c->n_calls + 1 > (int)c->starts.n)
    throw Error(LGBMHIP_UNSUPPORTED, "more quantized trees than the start-offset table holds (65536)");
  grow_trees(c, 1, false, 1.0);
  if (c->opts.use_quantized_grad) {
    hipLaunchKernelGGL(k_bump, dim3(1), dim3(1), 0, c->stream, c->call_ctr.p, 1);
    c->n_calls++;
  }
  HIP_CHECK(hipStreamSynchronize(c->stream));
  c->trees.push_back(build_tree(c, 0));
  c->pred_dirty = true;
}

// ---------------------------------------------------------------- prediction
// Flattened trees for prediction; categorical nodes carry the category
// bitset (raw values, tree.rs:399-418) or the bin bitset (binned rows,
// tree.rs:613-618).
static void build_pred(Ctx* c, bool binned) {
  std::vector<uint4> nodes;
  std::vector<int> noff, loff, nl;
  std::vector<double> lv;
  std::vector<unsigned> cats;
  for (const auto& t : c->trees) {
    noff.push_back((int)nodes.size());
    loff.push_back((int)lv.size());
    nl.push_back(t.num_leaves);
    if (t.num_leaves > 32768) throw Error(LGBMHIP_UNSUPPORTED, "trees above 32768 leaves in prediction");
    for (int i = 0; i < t.num_leaves - 1; ++i) {
      const int feat = binned ? t.split_feature_inner[i] : t.split_feature[i];
      if (feat < 0 || feat > 0xFFFF) throw Error(LGBMHIP_UNSUPPORTED, "split column above 65535 in prediction");
      const int mt = (t.decision_type[i] >> 2) & 3;
      const bool cat = t.decision_type[i] & 1;
      unsigned flags = (t.decision_type[i] & 2) ? 1u : 0u;
      flags |= mt == 1 ? 2u : mt == 2 ? 4u : 0u;
      flags |= cat ? 8u : 0u;
      uint4 n;
      n.x = (unsigned)feat | flags << 16 | (t.threshold_in_bin[i] & 0xFFu) << 24;
      n.y = ((unsigned)t.left_child[i] & 0xFFFFu) | ((unsigned)t.right_child[i] & 0xFFFFu) << 16;
      if (cat) {
        const int ci = (int)t.threshold_in_bin[i];
        const auto& bnd = binned ? t.cat_boundaries_inner : t.cat_boundaries;
        const auto& words = binned ? t.cat_threshold_inner : t.cat_threshold;
        n.z = (unsigned)cats.size();
        n.w = (unsigned)(bnd[ci + 1] - bnd[ci]);
        for (int w = bnd[ci]; w < bnd[ci + 1]; ++w) cats.push_back(words[w]);
      } else if (binned) {
        const auto& m = c->metas[t.split_feature_inner[i]];
        n.z = mt == 1 ? (unsigned)m.default_bin : (unsigned)(m.num_bin - 1);
        n.w = 0;
      } else {
        uint64_t bits;
        std::memcpy(&bits, &t.threshold[i], 8);
        n.z = (unsigned)bits;
        n.w = (unsigned)(bits >> 32);
      }
      nodes.push_back(n);
    }
    for (int l = 0; l < t.num_leaves; ++l) lv.push_back(t.leaf_value[l]);
  }
  c->p_T = (int)c->trees.size();
  c->p_nodes.alloc(std::max<size_t>(nodes.size(), 1));
  c->p_node_off.alloc(std::max<size_t>(noff.size(), 1));
  c->p_leaf_off.alloc(std::max<size_t>(loff.size(), 1));
  c->p_nleaves.alloc(std::max<size_t>(nl.size(), 1));
  c->p_leafval.alloc(std::max<size_t>(lv.size(), 1));
  c->p_cat.alloc(std::max<size_t>(cats.size(), 1));
  if (!cats.empty()) HIP_CHECK(hipMemcpy(c->p_cat.p, cats.data(), cats.size() * 4, hipMemcpyHostToDevice));
  if (!nodes.empty()) HIP_CHECK(hipMemcpy(c->p_nodes.p, nodes.data(), nodes.size() * sizeof(uint4), hipMemcpyHostToDevice));
  if (!noff.empty()) {
    HIP_CHECK(hipMemcpy(c->p_node_off.p, noff.data(), noff.size() * 4, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(c->p_leaf_off.p, loff.data(), loff.size() * 4, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(c->p_nleaves.p, nl.data(), nl.size() * 4, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(c->p_leafval.p, lv.data(), lv.size() * 8, hipMemcpyHostToDevice));
  }
  c->pred_dirty = false;
  c->p_binned = binned;
}

// Launches prediction over device-resident rows (raw f32 X or u8 bins) into
// a device [rows][K] output on the context's stream.
static void launch_predict(Ctx* c, const float* dX, const uint8_t* dbins, int64_t rows, int ncol, double* dout) {
  const bool binned = dbins != nullptr;
  if (binned && c->loaded_model) throw Error(LGBMHIP_UNSUPPORTED, "binned prediction of a model loaded from text");
  if (c->trees.empty()) throw Error(LGBMHIP_INVALID_PARAMETER, "no trees to predict with");
  if (c->pred_dirty || c->p_binned != binned) build_pred(c, binned);
  if (rows <= 0) return;
  // GBDT::InitPredict's iteration window (predict.rs:193-202)
  const int K = c->K;
  const int total = c->p_T / K;
  const int start = std::min(std::max(c->pred_start, 0), total);
  const int num = c->pred_num > 0 ? std::min(c->pred_num, total - start) : total - start;
  PredArgs a{};
  a.nodes = c->p_nodes.p;
  a.tree_node_off = c->p_node_off.p + (size_t)start * K;
  a.tree_leaf_off = c->p_leaf_off.p + (size_t)start * K;
  a.tree_num_leaves = c->p_nleaves.p + (size_t)start * K;
  a.leaf_val = c->p_leafval.p;
  a.cat_words = c->p_cat.p;
  a.T = num * K;
  a.K = K;
  a.rows = rows;
  a.ncol = binned ? c->F : ncol;
  a.X = dX;
  a.bins = dbins;
  a.out = dout;
  // odd word stride: lanes reading one column of their rows fall in distinct banks
  const int words = binned ? (a.ncol + 3) / 4 : a.ncol;
  const int sw = words | 1;
  a.stride = binned ? sw * 4 : sw;
  const size_t lds = (size_t)kPredRows * sw * 4;
  const dim3 grid((unsigned)((rows + kPredRows - 1) / kPredRows));
  // prediction_early_stop (predict.rs:175-190): binary, multiclass and ranking
  // objectives allow it; a model without an objective never stops early
  if (c->pred_es_freq > 0 && !(c->loaded_model && !c->loaded_head.has_objective)) {
    if (K > 16) throw Error(LGBMHIP_UNSUPPORTED, "early-stopped prediction of more than 16 classes");
    if (lds > (size_t)c->lds_max) throw Error(LGBMHIP_UNSUPPORTED, "rows too wide for early-stopped prediction");
    auto kern = binned ? k_predict_es<true> : k_predict_es<false>;
    HIP_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kern, grid, dim3(kPredRows), lds, c->stream, a, c->pred_es_freq, c->pred_es_margin);
    HIP_CHECK(hipGetLastError());
    return;
  }
  // waves sharing one staged 64-row tile: a raw f32 tile (~35 KB at 136
  // columns) allows few tiles per CU, so 16 waves share it; 4 for bin tiles
  // (1M x 136 rows, 1000 trees: 42 and 55 M rows/s)
  if (lds <= (size_t)c->lds_max) {
#ifndef LGBMHIP_PRED_BIN_W
#define LGBMHIP_PRED_BIN_W 4
#endif
#ifndef LGBMHIP_PRED_RAW_W
#define LGBMHIP_PRED_RAW_W 16
#endif
    auto kern = binned ? k_predict<true, true, LGBMHIP_PRED_BIN_W> : k_predict<false, true, LGBMHIP_PRED_RAW_W>;
    const int W = binned ? LGBMHIP_PRED_BIN_W : LGBMHIP_PRED_RAW_W;
    HIP_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kern, grid, dim3(kPredRows * W), lds, c->stream, a);
  } else {
    auto kern = binned ? k_predict<true, false, 1> : k_predict<false, false, 1>;
    hipLaunchKernelGGL(kern, grid, dim3(kPredRows), 0, c->stream, a);
  }
  HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------- validation sets and metrics
// Gbdt::add_valid (boosting.rs:378-411): bins checked and stored like the
// training bins, scores from init_score (or zeros), then every tree trained
// so far added one at a time in model order (tree.rs:635-652).
static int add_valid(Ctx* c, const uint8_t* bins, int64_t n, const float* label, const float* weight,
                     const int32_t* qb, int32_t nq, const double* init_score) {
  if (!c->N) throw Error(LGBMHIP_INVALID_DATA, "upload the training bins first");
  if (!c->opts_set) throw Error(LGBMHIP_INVALID_PARAMETER, "set the train options first");
  if (n <= 0 || n > INT32_MAX) throw Error(LGBMHIP_INVALID_DATA, "validation rows out of range");
  if (!bins || !label) throw Error(LGBMHIP_INVALID_PARAMETER, "validation bins and labels are required");
  set_device(c);
  const int K = c->opts.objective == 1 ? c->opts.num_class : 1;
  const int F = c->F;
  for (int64_t i = 0; i < n * F; ++i)
    if (bins[i] >= (uint8_t)c->metas[i % F].num_bin && c->metas[i % F].num_bin < 256)
      throw Error(LGBMHIP_INVALID_DATA, "validation bin value >= num_bin");
  if (qb) {
    if (nq <= 0 || qb[0] != 0 || qb[nq] != (int32_t)n) throw Error(LGBMHIP_INVALID_DATA, "query boundaries do not cover the validation data");
    for (int q = 0; q < nq; ++q)
      if (qb[q + 1] < qb[q]) throw Error(LGBMHIP_INVALID_DATA, "query boundaries must not decrease");
  }
  auto v = std::make_unique<ValidData>();
  v->n = n;
  std::vector<uint8_t> rm((size_t)n * c->S, 0);
  for (int64_t r = 0; r < n; ++r) pack_row(c, bins + (size_t)r * F, &rm[(size_t)r * c->S]);
  v->bins.alloc(rm.size());
  HIP_CHECK(hipMemcpy(v->bins.p, rm.data(), rm.size(), hipMemcpyHostToDevice));
  v->label.assign(label, label + n);
  v->d_label.alloc(n);
  HIP_CHECK(hipMemcpy(v->d_label.p, label, n * 4, hipMemcpyHostToDevice));
  v->has_weight = weight != nullptr;
  // Metric::new (metric.rs:238-247): the weight total in row order
  v->sum_weights = (double)n;
  if (weight) {
    v->d_weight.alloc(n);
    HIP_CHECK(hipMemcpy(v->d_weight.p, weight, n * 4, hipMemcpyHostToDevice));
    double sw = 0.0;
    for (int64_t i = 0; i < n; ++i) sw += (double)weight[i];
    v->sum_weights = sw;
  }
  if (qb) {
    v->qb.assign(qb, qb + nq + 1);
    v->d_qb.alloc(nq + 1);
    HIP_CHECK(hipMemcpy(v->d_qb.p, qb, (nq + 1) * 4, hipMemcpyHostToDevice));
  }
  const size_t cnt = (size_t)K * n;
  v->has_init = init_score != nullptr;
  if (init_score) {
    v->init.alloc(cnt);
    HIP_CHECK(hipMemcpy(v->init.p, init_score, cnt * 8, hipMemcpyHostToDevice));
  }
  const int saved_k = c->K;
  c->K = K;
  reset_valid_scores(c, *v);
  c->K = saved_k;
  if (c->state_ready && !c->trees.empty()) {
    if (c->pred_dirty || !c->p_binned) build_pred(c, true);
    const int words = (F + 3) / 4, sw = words | 1;
    const size_t lds = (size_t)kPredRows * sw * 4;
    if (lds > (size_t)c->lds_max) throw Error(LGBMHIP_UNSUPPORTED, "validation rows too wide to stage");
    auto kern = k_predict<true, true, LGBMHIP_PRED_BIN_W>;
    HIP_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const dim3 grid((unsigned)((n + kPredRows - 1) / kPredRows));
    // the prediction kernel reads rows of F bytes (the stored rows have stride S)
    DevBuf<uint8_t> packed;
    packed.alloc((size_t)n * F);
    HIP_CHECK(hipMemcpy(packed.p, bins, (size_t)n * F, hipMemcpyHostToDevice));
    for (int t = 0; t < c->p_T; ++t) {
      PredArgs a{};
      a.nodes = c->p_nodes.p;
      a.tree_node_off = c->p_node_off.p + t;
      a.tree_leaf_off = c->p_leaf_off.p + t;
      a.tree_num_leaves = c->p_nleaves.p + t;
      a.leaf_val = c->p_leafval.p;
      a.cat_words = c->p_cat.p;
      a.T = 1;
      a.K = 1;
      a.rows = n;
      a.ncol = F;
      a.stride = sw * 4;
      a.bins = packed.p;
      a.out = v->score.p + (size_t)(t % K) * n;
      a.accum = 1;
      hipLaunchKernelGGL(kern, grid, dim3(kPredRows * LGBMHIP_PRED_BIN_W), lds, c->stream, a);
    }
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(c->stream));
  }
  HIP_CHECK(hipStreamSynchronize(c->stream));
  c->valid.push_back(std::move(v));
  // captured trees hold the score walks of the sets they were captured with
  drop_graphs(c);
  return (int)c->valid.size();
}

// Metric ids of lgbmhip_eval.
enum { kMetricBinaryLogloss = 0, kMetricBinaryError = 1, kMetricAuc = 2, kMetricMultiLogloss = 3,
       kMetricMultiError = 4, kMetricNdcg = 5 };

// Metric::eval (metric.rs:525-545) on the scores of the training data
// (data 0) or validation set data.
static int eval_metric(Ctx* c, int data, const lgbmhip_metric_opts* o, double* out) {
  if (!c->state_ready) setup_training(c);
  set_device(c);
  if (data < 0 || data > (int)c->valid.size()) throw Error(LGBMHIP_INVALID_PARAMETER, "no such data set");
  const int K = c->K;
  const double* score;
  const float *label, *weight;
  const std::vector<float>* hlabel;
  const std::vector<int32_t>* hqb;
  const int* dqb = nullptr;
  int64_t n;
  double sum_weights;
  if (data == 0) {
    n = c->N;
    score = c->score.p;
    label = c->d_label.p;
    weight = c->has_weight ? c->d_weight.p : nullptr;
    hlabel = &c->label;
    hqb = &c->qb;
    sum_weights = c->sum_weights;
    if (!c->qb.empty()) {
      if (c->d_qb_eval.n != c->qb.size()) {
        c->d_qb_eval.alloc(c->qb.size());
        HIP_CHECK(hipMemcpy(c->d_qb_eval.p, c->qb.data(), c->qb.size() * 4, hipMemcpyHostToDevice));
      }
      dqb = c->d_qb_eval.p;
    }
  } else {
    const ValidData& v = *c->valid[data - 1];
    n = v.n;
    score = v.score.p;
    label = v.d_label.p;
    weight = v.has_weight ? v.d_weight.p : nullptr;
    hlabel = &v.label;
    hqb = &v.qb;
    dqb = v.qb.empty() ? nullptr : v.d_qb.p;
    sum_weights = v.sum_weights;
  }
  if (!c->mws) c->mws = metric_ws_new();
  const int obj = c->opts.objective;
  const double nle = -(double)std::log(1e-15f);   // consts.rs:12-14, the single-precision log
  switch (o->metric) {
    case kMetricBinaryLogloss:
    case kMetricBinaryError:
      if (obj != 0) throw Error(LGBMHIP_UNSUPPORTED, "binary metrics need the binary objective");
      out[0] = metric_point(c->mws, c->stream, o->metric == kMetricBinaryLogloss ? kBinaryLogloss : kBinaryError,
                            score, label, weight, n, 1, c->opts.sigmoid, 0, nle, sum_weights);
      return 1;
    case kMetricAuc:
      if (K != 1) throw Error(LGBMHIP_UNSUPPORTED, "AUC of a multiclass model");
      out[0] = metric_auc(c->mws, c->stream, score, label, weight, n, sum_weights);
      return 1;
    case kMetricMultiLogloss:
    case kMetricMultiError: {
      if (obj != 1) throw Error(LGBMHIP_UNSUPPORTED, "multiclass metrics need the multiclass objective");
      for (float l : *hlabel)
        if (!(l >= 0.0f && (int)l < K)) throw Error(LGBMHIP_INVALID_DATA, "multiclass label out of [0, num_class)");
      out[0] = metric_point(c->mws, c->stream, o->metric == kMetricMultiLogloss ? kMultiLogloss : kMultiError, score,
                            label, weight, n, K, 1.0, o->multi_error_top_k, nle, sum_weights);
      return 1;
    }
    case kMetricNdcg: {
      if (K != 1) throw Error(LGBMHIP_UNSUPPORTED, "NDCG of a multiclass model");
      if (!dqb || hqb->size() < 2) throw Error(LGBMHIP_INVALID_DATA, "The NDCG metric requires query information");
      if (weight) throw Error(LGBMHIP_UNSUPPORTED, "NDCG with row weights");
      // dcg.rs:27-52: default eval_at 1..5 and label gains 2^i - 1
      std::vector<int> at = o->num_eval_at > 0 ? std::vector<int>(o->eval_at, o->eval_at + o->num_eval_at)
                                               : std::vector<int>{1, 2, 3, 4, 5};
      std::vector<double> gain;
      if (o->num_label_gain > 0) gain.assign(o->label_gain, o->label_gain + o->num_label_gain);
      else {
        gain.push_back(0.0);
        for (int i = 1; i < 31; ++i) gain.push_back((double)((1 << i) - 1));
      }
      for (float l : *hlabel)
        if (!(l >= 0.0f) || (float)(int)l != l || (size_t)l >= gain.size())
          throw Error(LGBMHIP_INVALID_DATA, "NDCG labels must be integers below the label gain count");
      std::vector<double> disc(10000);
      for (int i = 0; i < 10000; ++i) disc[i] = 1.0 / std::log2(2.0 + (double)i);
      metric_ndcg(c->mws, c->stream, score, label, dqb, (int)hqb->size() - 1, at.data(), (int)at.size(), gain.data(),
                  (int)gain.size(), disc.data(), (int)disc.size(), o->num_threads, out);
      return (int)at.size();
    }
    default:
      throw Error(LGBMHIP_INVALID_PARAMETER, "unknown metric id");
  }
}

// ---------------------------------------------------------------- model text
// C printf's %.{p}g, as upstream writes model numbers (fmt.rs:6-44).
static std::string fmt_g(double v, int prec) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.*g", prec, v);
  return buf;
}
template <class T, class F>
static std::string join_fmt(const std::vector<T>& v, size_t n, F f) {
  std::string s;
  for (size_t i = 0; i < n && i < v.size(); ++i) {
    if (i) s += ' ';
    s += f(v[i]);
  }
  return s;
}

// Tree::to_model_string (tree.rs:718-770) of a non-linear tree.
static std::string tree_text(const HostTree& t) {
  const size_t n = (size_t)t.num_leaves, ni = n - 1;
  auto i2s = [](auto x) { return std::to_string((long long)x); };
  auto g17 = [](double x) { return fmt_g(x, 17); };
  auto g6 = [](double x) { return fmt_g(x, 6); };
  std::string s;
  s += "num_leaves=" + std::to_string(n) + "\n";
  s += "num_cat=" + std::to_string(t.num_cat) + "\n";
  s += "split_feature=" + join_fmt(t.split_feature, ni, i2s) + "\n";
  s += "split_gain=" + join_fmt(t.split_gain, ni, [&](float x) { return g6((double)x); }) + "\n";
  s += "threshold=" + join_fmt(t.threshold, ni, g17) + "\n";
  s += "decision_type=" + join_fmt(t.decision_type, ni, [](int8_t x) { return std::to_string((int)x); }) + "\n";
  s += "left_child=" + join_fmt(t.left_child, ni, i2s) + "\n";
  s += "right_child=" + join_fmt(t.right_child, ni, i2s) + "\n";
  s += "leaf_value=" + join_fmt(t.leaf_value, n, g17) + "\n";
  // at most the stored weights: a loaded one-leaf tree stores none (tree.rs:734-736)
  s += "leaf_weight=" + join_fmt(t.leaf_weight, n, g17) + "\n";
  s += "leaf_count=" + join_fmt(t.leaf_count, n, i2s) + "\n";
  s += "internal_value=" + join_fmt(t.internal_value, ni, g6) + "\n";
  s += "internal_weight=" + join_fmt(t.internal_weight, ni, g6) + "\n";
  s += "internal_count=" + join_fmt(t.internal_count, ni, i2s) + "\n";
  if (t.num_cat > 0) {
    s += "cat_boundaries=" + join_fmt(t.cat_boundaries, (size_t)t.num_cat + 1, i2s) + "\n";
    s += "cat_threshold=" + join_fmt(t.cat_threshold, t.cat_threshold.size(), i2s) + "\n";
  }
  s += "is_linear=0\n";
  s += "shrinkage=" + g6(t.shrinkage) + "\n\n";
  return s;
}

// GBDT::SaveModelToString (io.rs:13-88): header, tree blocks with their byte
// sizes, split-count importances (boosting.rs:1337-1360), parameters block.
static std::string model_text(const ModelHeader& h, const std::vector<HostTree>& trees, int num_iteration) {
  if (trees.empty()) throw Error(LGBMHIP_INVALID_PARAMETER, "no trees to save");
  std::vector<std::string> names;
  {
    std::string cur;
    for (const char* p = h.feature_names.c_str();; ++p) {
      if (*p == ' ' || *p == 0) {
        names.push_back(cur);
        cur.clear();
        if (*p == 0) break;
      } else {
        cur += *p;
      }
    }
  }
  if ((int)names.size() != h.ncol) throw Error(LGBMHIP_INVALID_PARAMETER, "one feature name per column");
  std::string s = "tree\nversion=v4\n";
  s += "num_class=" + std::to_string(h.num_class) + "\n";
  s += "num_tree_per_iteration=" + std::to_string(h.ntpi) + "\n";
  s += "label_index=" + std::to_string(h.label_index) + "\n";
  s += "max_feature_idx=" + std::to_string(h.ncol - 1) + "\n";
  if (h.has_objective) s += "objective=" + h.objective + "\n";
  s += "feature_names=" + h.feature_names + "\n";
  s += "feature_infos=" + h.feature_infos + "\n";
  size_t used = trees.size();
  if (num_iteration > 0) used = std::min(used, (size_t)num_iteration * h.ntpi);
  std::vector<std::string> ts;
  std::string sizes;
  for (size_t i = 0; i < used; ++i) {
    ts.push_back("Tree=" + std::to_string(i) + "\n" + tree_text(trees[i]) + "\n");
    sizes += (i ? " " : "") + std::to_string(ts.back().size());
  }
  s += "tree_sizes=" + sizes + "\n\n";
  for (auto& t : ts) s += t;
  s += "end of trees\n";
  std::vector<double> imp(h.ncol, 0.0);
  for (size_t i = 0; i < used; ++i) {
    const HostTree& t = trees[i];
    for (int k = 0; k + 1 < t.num_leaves; ++k)
      if (t.split_gain[k] > 0.0f && t.split_feature[k] >= 0 && t.split_feature[k] < h.ncol) imp[t.split_feature[k]] += 1.0;
  }
  std::vector<std::pair<size_t, int>> pairs;
  for (int f = 0; f < h.ncol; ++f)
    if ((size_t)imp[f] > 0) pairs.emplace_back((size_t)imp[f], f);
  std::stable_sort(pairs.begin(), pairs.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
  s += "\nfeature_importances:\n";
  for (auto& pr : pairs) s += names[pr.second] + "=" + std::to_string(pr.first) + "\n";
  if (h.has_parameters) s += "\nparameters:\n" + h.parameters + "\nend of parameters\n";
  return s;
}

// A model text's trees and header.
struct ParsedModel {
  std::vector<HostTree> trees;
  ModelHeader head;
  int objective = 0;   // 0 binary, 1 multiclass, 2 lambdarank
};

// GBDT::LoadModelFromString (io.rs:156-343) and Tree::from_model_str
// (tree.rs:869-1047): header fields, every tree field the writer writes
// (exactly the expected count of values each), the parameters block; the
// checks the reference makes, plus a tree-shape check (each node but the
// root, and each leaf, referenced exactly once). No device work.
static ParsedModel parse_model(const char* text) {
  if (!text) throw Error(LGBMHIP_INVALID_PARAMETER, "no model text");
  std::vector<std::string> lines;
  for (const char* p = text;;) {
    const char* e = std::strchr(p, '\n');
    std::string ln = e ? std::string(p, e - p) : std::string(p);
    if (!ln.empty() && ln.back() == '\r') ln.pop_back();
    lines.push_back(std::move(ln));
    if (!e) break;
    p = e + 1;
  }
  auto key_of = [](const std::string& ln) { return ln.substr(0, ln.find('=')); };
  auto val_of = [](const std::string& ln) {
    const size_t eq = ln.find('=');
    return eq == std::string::npos ? std::string() : ln.substr(eq + 1);
  };
  std::unordered_map<std::string, std::string> hk;
  size_t i = 0;
  for (; i < lines.size() && lines[i].rfind("Tree=", 0) != 0; ++i) {
    if (lines[i].empty()) continue;
    if (lines[i] == "end of trees") break;
    hk[key_of(lines[i])] = val_of(lines[i]);
  }
  auto head_int = [&](const char* k, long long lo, long long hi) {
    long long v;
    if (!hk.count(k)) throw Error(LGBMHIP_MODEL_FORMAT, std::string("Model file doesn't specify ") + k);
    if (!parse_int(hk[k], lo, hi, &v)) throw Error(LGBMHIP_MODEL_FORMAT, std::string("bad ") + k);
    return (int)v;
  };
  ParsedModel m;
  ModelHeader& h = m.head;
  h.num_class = head_int("num_class", 1, 1 << 20);
  h.ntpi = hk.count("num_tree_per_iteration") ? head_int("num_tree_per_iteration", 1, 1 << 20) : h.num_class;
  h.label_index = head_int("label_index", -(1LL << 31), (1LL << 31) - 1);
  h.ncol = head_int("max_feature_idx", 0, (1LL << 31) - 2) + 1;
  if (!hk.count("feature_names")) throw Error(LGBMHIP_MODEL_FORMAT, "Model file doesn't contain feature_names");
  if (!hk.count("feature_infos")) throw Error(LGBMHIP_MODEL_FORMAT, "Model file doesn't contain feature_infos");
  h.feature_names = hk["feature_names"];
  h.feature_infos = hk["feature_infos"];
  auto count_fields = [](const std::string& s) { return (int)std::count(s.begin(), s.end(), ' ') + 1; };
  if (count_fields(h.feature_names) != h.ncol) throw Error(LGBMHIP_MODEL_FORMAT, "Wrong size of feature_names");
  if (count_fields(h.feature_infos) != h.ncol) throw Error(LGBMHIP_MODEL_FORMAT, "Wrong size of feature_infos");
  h.has_objective = hk.count("objective") > 0;
  h.objective = h.has_objective ? hk["objective"] : std::string();
  // The prediction here is the raw sum of binary, multiclass and lambdarank
  // models, with their early-stopping margins (predict.rs:152-190), or of a
  // model without an objective (a custom one), which never stops early.
  // Forests (average_output) and other objectives are not served.
  if (hk.count("average_output")) throw Error(LGBMHIP_UNSUPPORTED, "models with average_output (random forest)");
  const std::string& obj = h.objective;
  const bool obj_binary = obj.rfind("binary", 0) == 0, obj_multi = obj.rfind("multiclass", 0) == 0,
             obj_rank = obj.rfind("lambdarank", 0) == 0;
  if (h.has_objective && !obj_binary && !obj_multi && !obj_rank)
    throw Error(LGBMHIP_UNSUPPORTED, "model objective '" + obj + "' (binary, multiclass and lambdarank are served)");
  m.objective = obj_multi ? 1 : obj_rank ? 2 : h.has_objective ? 0 : h.ntpi > 1 ? 1 : 0;

  while (i < lines.size() && lines[i].rfind("Tree=", 0) == 0) {
    HostTree t;
    std::unordered_map<std::string, std::string> kv;
    for (++i; i < lines.size() && !lines[i].empty() && lines[i].rfind("Tree=", 0) != 0 && lines[i] != "end of trees";
         ++i) {
      if (lines[i].find('=') == std::string::npos) throw Error(LGBMHIP_MODEL_FORMAT, "bad tree line: " + lines[i]);
      kv[key_of(lines[i])] = val_of(lines[i]);
    }
    while (i < lines.size() && lines[i].empty()) ++i;
    auto need = [&](const char* k) -> const std::string& {
      if (!kv.count(k))
        throw Error(LGBMHIP_MODEL_FORMAT, std::string("Tree model string format error, should contain ") + k + " field");
      return kv[k];
    };
    auto ints = [&](const std::string& s, size_t n, const char* name, long long lo, long long hi) {
      const auto toks = split_ws(s);
      if (toks.size() != n) throw Error(LGBMHIP_MODEL_FORMAT, std::string(name) + " has a wrong number of values");
      std::vector<long long> v(n);
      for (size_t k = 0; k < n; ++k)
        if (!parse_int(toks[k], lo, hi, &v[k])) throw Error(LGBMHIP_MODEL_FORMAT, std::string("bad value in ") + name);
      return v;
    };
    auto reals = [&](const std::string& s, size_t n, const char* name, bool legacy) {
      const auto toks = split_ws(s);
      if (toks.size() != n) throw Error(LGBMHIP_MODEL_FORMAT, std::string(name) + " has a wrong number of values");
      std::vector<double> v(n);
      for (size_t k = 0; k < n; ++k)
        if (!(legacy ? atof_legacy(toks[k], &v[k]) : parse_f64(toks[k], &v[k])))
          throw Error(LGBMHIP_MODEL_FORMAT, std::string("bad value in ") + name);
      return v;
    };
    const long long I32MAX = (1LL << 31) - 1, I32MIN = -(1LL << 31);
    const int n = (int)ints(need("num_leaves"), 1, "num_leaves", 1, 1 << 24)[0];
    t.num_cat = (int)ints(need("num_cat"), 1, "num_cat", 0, I32MAX)[0];
    if (kv.count("is_linear") && kv["is_linear"] != "0") throw Error(LGBMHIP_UNSUPPORTED, "linear trees");
    const int ni = n - 1;
    t.num_leaves = n;
    t.leaf_value = reals(need("leaf_value"), n, "leaf_value", false);
    t.shrinkage = 1.0;
    if (kv.count("shrinkage")) {
      double sh;
      if (atof_legacy(kv["shrinkage"], &sh)) t.shrinkage = sh;
    }
    t.leaf_count.assign(n, 0);
    if (kv.count("leaf_count")) {
      const auto lc = ints(kv["leaf_count"], n, "leaf_count", I32MIN, I32MAX);
      for (int k = 0; k < n; ++k) t.leaf_count[k] = (int)lc[k];
    }
    t.leaf_parent.assign(n, -1);
    t.leaf_depth.assign(n, 0);
    t.leaf_weight.clear();   // a one-leaf tree stores no weights (tree.rs:962-963)
    if (ni > 0) {
      t.leaf_weight = kv.count("leaf_weight") ? reals(kv["leaf_weight"], n, "leaf_weight", false)
                                              : std::vector<double>(n, 0.0);
      const auto lc = ints(need("left_child"), ni, "left_child", I32MIN, I32MAX);
      const auto rc = ints(need("right_child"), ni, "right_child", I32MIN, I32MAX);
      const auto sf = ints(need("split_feature"), ni, "split_feature", I32MIN, I32MAX);
      const auto th = reals(need("threshold"), ni, "threshold", false);
      const auto sg = kv.count("split_gain") ? reals(kv["split_gain"], ni, "split_gain", true) : std::vector<double>(ni, 0.0);
      const auto ic = kv.count("internal_count") ? ints(kv["internal_count"], ni, "internal_count", I32MIN, I32MAX)
                                                 : std::vector<long long>(ni, 0);
      t.internal_value = kv.count("internal_value") ? reals(kv["internal_value"], ni, "internal_value", true)
                                                    : std::vector<double>(ni, 0.0);
      t.internal_weight = kv.count("internal_weight") ? reals(kv["internal_weight"], ni, "internal_weight", true)
                                                      : std::vector<double>(ni, 0.0);
      const auto dt = kv.count("decision_type") ? ints(kv["decision_type"], ni, "decision_type", -128, 127)
                                                : std::vector<long long>(ni, 0);
      if (t.num_cat > 0) {
        const size_t nb = (size_t)t.num_cat + 1;
        if (!kv.count("cat_boundaries")) throw Error(LGBMHIP_MODEL_FORMAT, "Tree model should contain cat_boundaries field.");
        const auto cb = ints(kv["cat_boundaries"], nb, "cat_boundaries", I32MIN, I32MAX);
        // tree.rs:993-1009: boundaries ascend from 0 and end at the bitset words' count
        if (cb[0] != 0) throw Error(LGBMHIP_MODEL_FORMAT, "cat_boundaries must be non-decreasing from 0");
        for (size_t k = 1; k < nb; ++k)
          if (cb[k] < cb[k - 1]) throw Error(LGBMHIP_MODEL_FORMAT, "cat_boundaries must be non-decreasing from 0");
        if (!kv.count("cat_threshold")) throw Error(LGBMHIP_MODEL_FORMAT, "Tree model should contain cat_threshold field");
        const auto ct = ints(kv["cat_threshold"], (size_t)cb[nb - 1], "cat_threshold", 0, 4294967295LL);
        t.cat_boundaries.assign(cb.begin(), cb.end());
        t.cat_threshold.assign(ct.begin(), ct.end());
      }
      for (int k = 0; k < ni; ++k) {
        // a categorical node's threshold is its bitset's index (tree.rs:400-403, 1011-1018);
        // prediction reads that index from threshold_in_bin
        const bool cat = (dt[k] & 1) != 0;
        if (cat && !(th[k] >= 0.0 && (int)th[k] < t.num_cat))
          throw Error(LGBMHIP_MODEL_FORMAT, "categorical split index out of range");
        for (long long ch : {lc[k], rc[k]})
          if (ch >= 0 ? ch >= ni : ~ch >= n) throw Error(LGBMHIP_MODEL_FORMAT, "child index out of range");
        if (sf[k] < 0 || sf[k] >= h.ncol) throw Error(LGBMHIP_MODEL_FORMAT, "split_feature out of range");
        t.split_feature.push_back((int)sf[k]);
        t.split_feature_inner.push_back((int)sf[k]);
        t.threshold.push_back(th[k]);
        t.threshold_in_bin.push_back(cat ? (uint32_t)th[k] : 0u);
        t.decision_type.push_back((int8_t)dt[k]);
        t.left_child.push_back((int)lc[k]);
        t.right_child.push_back((int)rc[k]);
        t.split_gain.push_back((float)sg[k]);
        t.internal_count.push_back((int)ic[k]);
      }
      // Every internal node but the root, and every leaf, is some node's child
      // exactly once. A walk from the root then meets each node at most once,
      // so the device's node-chain loops end within num_leaves - 1 steps.
      std::vector<int> refs((size_t)(ni + n), 0);
      for (int k = 0; k < ni; ++k)
        for (int ch : {t.left_child[k], t.right_child[k]}) ++refs[ch >= 0 ? (size_t)ch : (size_t)(ni + ~ch)];
      if (refs[0] != 0) throw Error(LGBMHIP_MODEL_FORMAT, "the root is a child");
      for (size_t r = 1; r < refs.size(); ++r)
        if (refs[r] != 1) throw Error(LGBMHIP_MODEL_FORMAT, "tree node referenced other than once");
    }
    m.trees.push_back(std::move(t));
  }
  if (m.trees.empty()) throw Error(LGBMHIP_MODEL_FORMAT, "model text without trees");
  // the parameters block after the trees (io.rs:272-290): its lines, blank lines skipped
  for (; i < lines.size(); ++i) {
    if (lines[i] == "parameters:") {
      h.has_parameters = true;
      for (++i; i < lines.size() && lines[i] != "end of parameters"; ++i)
        if (!lines[i].empty()) h.parameters += lines[i] + "\n";
      break;
    }
  }
  return m;
}

static void load_model(Ctx* c, const char* text) {
  ParsedModel m = parse_model(text);
  c->trees = std::move(m.trees);
  c->K = m.head.ntpi;
  c->ncol = m.head.ncol;
  c->opts.objective = m.objective;
  c->opts.num_class = m.head.num_class;
  c->loaded_head = m.head;
  c->opts_set = true;
  c->pred_dirty = true;
  c->state_ready = false;
  c->loaded_model = true;
}

// The header of the context's model: the loaded text's, or the trained
// booster's (objective to_model_string: binary.rs:168-170,
// multiclass.rs:169-171, rank.rs:338-340). names / infos / parameters
// replace the header's when given.
static std::string save_model(Ctx* c, int num_iteration, const char* names, const char* infos, const char* params) {
  ModelHeader h;
  if (c->loaded_model) {
    h = c->loaded_head;
  } else {
    h.ncol = c->ncol;
    h.ntpi = c->K;
    h.num_class = c->opts.objective == 1 ? c->opts.num_class : 1;
    if (c->opts.objective == 0) h.objective = "binary sigmoid:" + fmt_g(c->opts.sigmoid, 6);
    else if (c->opts.objective == 1) h.objective = "multiclass num_class:" + std::to_string(c->opts.num_class);
    else h.objective = "lambdarank";
    for (int f = 0; f < h.ncol; ++f) {
      h.feature_names += (f ? " Column_" : "Column_") + std::to_string(f);
      h.feature_infos += f ? " none" : "none";
    }
  }
  if (names) h.feature_names = names;
  if (infos) h.feature_infos = infos;
  if (params) {
    h.has_parameters = true;
    h.parameters = params;
    if (!h.parameters.empty() && h.parameters.back() != '\n') h.parameters += '\n';
  }
  return model_text(h, c->trees, num_iteration);
}

// PredictKind::LeafIndex over host raw rows: out [rows][trees of the window].
static int64_t predict_leaf(Ctx* c, const float* X, int64_t rows, int ncol, double* out) {
  set_device(c);
  if (c->trees.empty()) throw Error(LGBMHIP_INVALID_PARAMETER, "no trees to predict with");
  if (c->pred_dirty || c->p_binned) build_pred(c, false);
  const int K = c->K, total = c->p_T / K;
  const int start = std::min(std::max(c->pred_start, 0), total);
  const int num = c->pred_num > 0 ? std::min(c->pred_num, total - start) : total - start;
  const int T = num * K;
  if (rows <= 0 || T == 0) return T;
  PredArgs a{};
  a.nodes = c->p_nodes.p;
  a.tree_node_off = c->p_node_off.p + (size_t)start * K;
  a.tree_leaf_off = c->p_leaf_off.p + (size_t)start * K;
  a.tree_num_leaves = c->p_nleaves.p + (size_t)start * K;
  a.leaf_val = c->p_leafval.p;
  a.cat_words = c->p_cat.p;
  a.T = T;
  a.K = K;
  a.rows = rows;
  a.ncol = ncol;
  const int sw = ncol | 1;
  a.stride = sw;
  const size_t lds = (size_t)kPredRows * sw * 4;
  if (lds > (size_t)c->lds_max) throw Error(LGBMHIP_UNSUPPORTED, "rows too wide for leaf prediction");
  DevBuf<float> dx;
  DevBuf<double> dout;
  dx.alloc((size_t)rows * ncol);
  dout.alloc((size_t)rows * T);
  HIP_CHECK(hipMemcpyAsync(dx.p, X, (size_t)rows * ncol * 4, hipMemcpyHostToDevice, c->stream));
  a.X = dx.p;
  a.out = dout.p;
  HIP_CHECK(hipFuncSetAttribute((const void*)k_predict_leaf<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds));
  hipLaunchKernelGGL(k_predict_leaf<false>, dim3((unsigned)((rows + kPredRows - 1) / kPredRows)), dim3(kPredRows),
                     lds, c->stream, a);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(out, dout.p, (size_t)rows * T * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  return T;
}

// Gbdt::rollback_one_iter (boosting.rs:1215-1239): the last iteration's trees,
// shrunk by -1, added to the training and validation scores (binned walks),
// then dropped.
static void rollback_one_iter(Ctx* c) {
  if (!c->state_ready || c->iter == 0 || (int)c->trees.size() < c->K) return;
  set_device(c);
  const int K = c->K;
  if (c->pred_dirty || !c->p_binned) build_pred(c, true);
  const int first = c->p_T - K;
  // leaf values of the K trees as shrink(-1) leaves them (tree.rs:303-319)
  std::vector<int> loff(c->p_T);
  HIP_CHECK(hipMemcpy(loff.data(), c->p_leaf_off.p, c->p_T * 4, hipMemcpyDeviceToHost));
  DevBuf<double> neg;
  neg.alloc(c->p_leafval.n);
  HIP_CHECK(hipMemcpy(neg.p, c->p_leafval.p, c->p_leafval.n * 8, hipMemcpyDeviceToDevice));
  for (int k = 0; k < K; ++k) {
    const HostTree& t = c->trees[(size_t)first + k];
    std::vector<double> v(t.num_leaves);
    for (int l = 0; l < t.num_leaves; ++l) v[l] = HostTree::r0(t.leaf_value[l] * -1.0);
    HIP_CHECK(hipMemcpy(neg.p + loff[first + k], v.data(), v.size() * 8, hipMemcpyHostToDevice));
  }
  auto add_tree = [&](const uint8_t* bins, int64_t n, int stride_cols, double* score, int t) {
    const int words = (stride_cols + 3) / 4, sw = words | 1;
    const size_t lds = (size_t)kPredRows * sw * 4;
    auto kern = k_predict<true, true, LGBMHIP_PRED_BIN_W>;
    HIP_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    PredArgs a{};
    a.nodes = c->p_nodes.p;
    a.tree_node_off = c->p_node_off.p + t;
    a.tree_leaf_off = c->p_leaf_off.p + t;
    a.tree_num_leaves = c->p_nleaves.p + t;
    a.leaf_val = neg.p;
    a.cat_words = c->p_cat.p;
    a.T = 1;
    a.K = 1;
    a.rows = n;
    a.ncol = stride_cols;   // rows of S bytes: the padding columns are never read by a node
    a.stride = sw * 4;
    a.bins = bins;
    a.out = score;
    a.accum = 1;
    hipLaunchKernelGGL(kern, dim3((unsigned)((n + kPredRows - 1) / kPredRows)), dim3(kPredRows * LGBMHIP_PRED_BIN_W),
                       lds, c->stream, a);
    HIP_CHECK(hipGetLastError());
  };
  // packed rows with 4-bit fields: prediction reads one u8 bin per feature,
  // so each set's rows are unpacked first (one u8 bin per feature)
  DevBuf<uint8_t> unpacked;
  auto rows_of = [&](const uint8_t* bins, int64_t n, int& cols) -> const uint8_t* {
    cols = c->S;
    if (!c->has_u4) return bins;
    unpacked.alloc((size_t)n * c->F);
    hipLaunchKernelGGL(k_unpack_rows, dim3(grid_for(n * c->F, 256, c->cus * 16)), dim3(256), 0, c->stream, bins,
                       (long long)n, c->S, c->F, c->d_feats.p, unpacked.p);
    HIP_CHECK(hipGetLastError());
    cols = c->F;
    return unpacked.p;
  };
  for (int k = 0; k < K; ++k) {
    int cols = 0;
    const uint8_t* rb = rows_of(c->bins_rm.p, c->N, cols);
    add_tree(rb, c->N, cols, c->score.p + (size_t)k * c->N, first + k);
    for (auto& v : c->valid) {
      const uint8_t* vb = rows_of(v->bins.p, v->n, cols);
      add_tree(vb, v->n, cols, v->score.p + (size_t)k * v->n, first + k);
    }
  }
  HIP_CHECK(hipStreamSynchronize(c->stream));
  c->trees.resize(c->trees.size() - K);
  c->iter--;
  c->pred_dirty = true;
}

static void predict(Ctx* c, const float* X, const uint8_t* bins, int64_t rows, int ncol, double* out) {
  set_device(c);
  DevBuf<float> dx;
  DevBuf<uint8_t> db;
  DevBuf<double> dout;
  dout.alloc(std::max<size_t>((size_t)rows * c->K, 1));
  if (X) {
    dx.alloc(std::max<size_t>((size_t)rows * ncol, 1));
    HIP_CHECK(hipMemcpyAsync(dx.p, X, (size_t)rows * ncol * 4, hipMemcpyHostToDevice, c->stream));
  } else {
    db.alloc(std::max<size_t>((size_t)rows * c->F, 1));
    HIP_CHECK(hipMemcpyAsync(db.p, bins, (size_t)rows * c->F, hipMemcpyHostToDevice, c->stream));
  }
  launch_predict(c, X ? dx.p : nullptr, X ? nullptr : db.p, rows, ncol, dout.p);
  HIP_CHECK(hipMemcpyAsync(out, dout.p, (size_t)rows * c->K * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
}

}  // namespace lgbmhip

// ==================================================================== C ABI
using namespace lgbmhip;

// The no-allocation variant's work buffer: one per device (each GPU holds its
// own copy at its own address), shared by every call on that device.
static __device__ __align__(256) unsigned scratch[kNoAllocScratchWords];
static std::mutex g_scratch_mu;
static std::vector<unsigned*> g_scratch_ptr;   // per device: scratch's device address, resolved once

// Resolves scratch's address on `device` (current) once; host code then uses
// that pointer, never the array's name (a host-side stand-in).
static unsigned* scratch_of(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize((size_t)device + 1, nullptr);
  if (!g_scratch_ptr[(size_t)device]) {
    void* ptr = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw Error(LGBMHIP_INTERNAL, "no device address for the partition scratch");
    g_scratch_ptr[(size_t)device] = (unsigned*)ptr;
  }
  return g_scratch_ptr[(size_t)device];
}

struct lgbmhip_ctx {
  Ctx c;
};
struct lgbmhip_bins {
  std::unique_ptr<Binned> b;
};

static thread_local std::string g_err;

template <class F>
static lgbmhip_status guarded(lgbmhip_ctx* ctx, F&& fn) {
  try {
    fn();
    return LGBMHIP_OK;
  } catch (const Error& e) {
    if (ctx) ctx->c.err = e.what();
    g_err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    if (ctx) ctx->c.err = std::string("internal error: ") + e.what();
    g_err = std::string("internal error: ") + e.what();
    return LGBMHIP_INTERNAL;
  } catch (...) {
    if (ctx) ctx->c.err = "internal error";
    g_err = "internal error";
    return LGBMHIP_INTERNAL;
  }
}

extern "C" {

const char* lgbmhip_last_error(const lgbmhip_ctx* ctx) { return ctx ? ctx->c.err.c_str() : g_err.c_str(); }
const char* lgbmhip_global_last_error(void) { return g_err.c_str(); }

lgbmhip_status lgbmhip_bin_dense(const float* X, int64_t n, int32_t ncol, const int32_t* categorical, int32_t n_cat,
                                 const lgbmhip_bin_params* params, lgbmhip_bins** out) {
  return guarded(nullptr, [&] {
    if (!X || !params || !out) throw Error(LGBMHIP_INVALID_PARAMETER, "null argument");
    auto* b = new lgbmhip_bins();
    try {
      b->b.reset(bin_dense(X, n, ncol, categorical, n_cat, *params));
    } catch (...) {
      delete b;
      throw;
    }
    *out = b;
  });
}
lgbmhip_status lgbmhip_find_mappers(const float* X, int64_t n, int32_t ncol, const int32_t* categorical,
                                    int32_t n_cat, const lgbmhip_bin_params* params, lgbmhip_bins** out) {
  return guarded(nullptr, [&] {
    if (!X || !params || !out) throw Error(LGBMHIP_INVALID_PARAMETER, "null argument");
    auto* b = new lgbmhip_bins();
    try {
      b->b.reset(find_mappers(X, n, ncol, categorical, n_cat, *params));
    } catch (...) {
      delete b;
      throw;
    }
    *out = b;
  });
}
void lgbmhip_bins_free(lgbmhip_bins* b) { delete b; }
int32_t lgbmhip_bins_num_features(const lgbmhip_bins* b) { return b ? (int32_t)b->b->used.size() : 0; }
int32_t lgbmhip_bins_num_groups(const lgbmhip_bins* b) { return b ? (int32_t)b->b->num_groups : 0; }
int64_t lgbmhip_bins_num_upper(const lgbmhip_bins* b) {
  int64_t s = 0;
  for (int c : b->b->used) s += b->b->mappers[c].num_bin;
  return s;
}
lgbmhip_status lgbmhip_bins_metas(const lgbmhip_bins* b, lgbmhip_feature_meta* metas, double* upper) {
  return guarded(nullptr, [&] {
    size_t off = 0;
    for (size_t f = 0; f < b->b->used.size(); ++f) {
      const Mapper& m = b->b->mappers[b->b->used[f]];
      metas[f] = lgbmhip_feature_meta{b->b->used[f], m.num_bin, m.missing_type, (int32_t)m.default_bin,
                                      (int32_t)m.most_freq_bin, m.bin_type};
      for (int i = 0; i < m.num_bin; ++i) upper[off + i] = m.bin_type == 1 ? (double)m.categories[i] : m.bounds[i];
      off += (size_t)m.num_bin;
    }
  });
}
lgbmhip_status lgbmhip_bins_matrix(const lgbmhip_bins* b, uint8_t* out) {
  return guarded(nullptr, [&] {
    if (b->b->bins.size() != (size_t)b->b->n * b->b->used.size())
      throw Error(LGBMHIP_INVALID_PARAMETER, "these mappers hold no bin matrix (rows are binned on the device)");
    std::memcpy(out, b->b->bins.data(), b->b->bins.size());
  });
}
lgbmhip_status lgbmhip_bins_apply(const lgbmhip_bins* b, const float* X, int64_t n, int32_t ncol, uint8_t* out) {
  return guarded(nullptr, [&] { apply_bins(*b->b, X, n, ncol, out); });
}

lgbmhip_status lgbmhip_create(int32_t device, lgbmhip_ctx** out) {
  return guarded(nullptr, [&] {
    int n = 0;
    HIP_CHECK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) throw Error(LGBMHIP_INVALID_PARAMETER, "no such HIP device");
    auto* x = new lgbmhip_ctx();
    x->c.device = device;
    try {
      HIP_CHECK(hipSetDevice(device));
      HIP_CHECK(hipStreamCreateWithFlags(&x->c.stream, hipStreamNonBlocking));
      hipDeviceProp_t prop;
      HIP_CHECK(hipGetDeviceProperties(&prop, device));
      x->c.cus = prop.multiProcessorCount;
      scratch_of(device);   // the no-allocation entry point's work buffer on this device
      x->c.lds_max = prop.sharedMemPerBlock;
      if (prop.maxSharedMemoryPerMultiProcessor > 0)
        x->c.lds_max = std::min<size_t>(std::max<size_t>(x->c.lds_max, 65536), prop.maxSharedMemoryPerMultiProcessor);
    } catch (...) {
      delete x;
      throw;
    }
    *out = x;
  });
}
void lgbmhip_destroy(lgbmhip_ctx* ctx) { delete ctx; }

lgbmhip_status lgbmhip_upload_bins(lgbmhip_ctx* ctx, const lgbmhip_feature_meta* metas, int32_t num_features,
                                   const double* upper, const uint8_t* bins, int64_t num_rows,
                                   int32_t num_total_features) {
  return guarded(ctx, [&] { upload_bins(&ctx->c, metas, num_features, upper, bins, num_rows, num_total_features); });
}

lgbmhip_status lgbmhip_set_metadata(lgbmhip_ctx* ctx, const float* label, const float* weight,
                                    const int32_t* query_boundaries, int32_t num_queries) {
  return guarded(ctx, [&] {
    Ctx* c = &ctx->c;
    if (!c->N) throw Error(LGBMHIP_INVALID_DATA, "upload bins first");
    set_device(c);
    c->label.assign(label, label + c->N);
    c->d_label.alloc(c->N);
    HIP_CHECK(hipMemcpy(c->d_label.p, label, c->N * 4, hipMemcpyHostToDevice));
    c->has_weight = weight != nullptr;
    c->sum_weights = (double)c->N;
    if (weight) {
      c->d_weight.alloc(c->N);
      HIP_CHECK(hipMemcpy(c->d_weight.p, weight, c->N * 4, hipMemcpyHostToDevice));
      double sw = 0.0;
      for (int64_t i = 0; i < c->N; ++i) sw += (double)weight[i];
      c->sum_weights = sw;
    }
    if (query_boundaries) c->qb.assign(query_boundaries, query_boundaries + num_queries + 1);
    c->state_ready = false;
  });
}

lgbmhip_status lgbmhip_set_split_params(lgbmhip_ctx* ctx, const lgbmhip_split_params* p) {
  return guarded(ctx, [&] {
    if (p->num_leaves < 2 || p->num_leaves > 131072) throw Error(LGBMHIP_INVALID_PARAMETER, "num_leaves out of range");
    if (p->min_data_in_leaf < 0) throw Error(LGBMHIP_INVALID_PARAMETER, "min_data_in_leaf < 0");
    SplitParamsDev d{p->lambda_l1, p->lambda_l2, p->max_delta_step, p->path_smooth, p->min_sum_hessian_in_leaf,
                     p->min_gain_to_split, p->cat_l2, p->cat_smooth, p->min_data_in_leaf, p->max_cat_to_onehot,
                     p->max_cat_threshold, p->min_data_per_group, p->num_leaves, p->max_depth};
    ctx->c.sp = d;
    ctx->c.sp_set = true;
    ctx->c.state_ready = false;
  });
}

lgbmhip_status lgbmhip_set_train_opts(lgbmhip_ctx* ctx, const lgbmhip_train_opts* o) {
  return guarded(ctx, [&] {
    if (o->objective < 0 || o->objective > 2) throw Error(LGBMHIP_INVALID_PARAMETER, "unknown objective");
    if (o->sigmoid <= 0.0) throw Error(LGBMHIP_INVALID_PARAMETER, "sigmoid must be > 0");
    ctx->c.opts = *o;
    ctx->c.opts_set = true;
    ctx->c.state_ready = false;
  });
}

lgbmhip_status lgbmhip_train_iteration(lgbmhip_ctx* ctx, int32_t* finished) {
  return guarded(ctx, [&] { *finished = train_iteration(&ctx->c); });
}
lgbmhip_status lgbmhip_train_tree(lgbmhip_ctx* ctx, const float* grad, const float* hess) {
  return guarded(ctx, [&] { train_tree(&ctx->c, grad, hess); });
}
lgbmhip_status lgbmhip_sync(lgbmhip_ctx* ctx) {
  return guarded(ctx, [&] { HIP_CHECK(hipStreamSynchronize(ctx->c.stream)); });
}
int32_t lgbmhip_num_trees(lgbmhip_ctx* ctx) { return (int32_t)ctx->c.trees.size(); }
int32_t lgbmhip_tree_num_leaves(lgbmhip_ctx* ctx, int32_t t) {
  if (t < 0 || t >= (int32_t)ctx->c.trees.size()) return -1;
  return ctx->c.trees[t].num_leaves;
}
lgbmhip_status lgbmhip_tree_arrays(lgbmhip_ctx* ctx, int32_t t, int32_t* split_feature, uint32_t* threshold_in_bin,
                                   double* threshold, int8_t* decision_type, int32_t* left_child, int32_t* right_child,
                                   float* split_gain, double* leaf_value, int32_t* leaf_count, double* leaf_weight,
                                   double* internal_value) {
  return guarded(ctx, [&] {
    if (t < 0 || t >= (int32_t)ctx->c.trees.size()) throw Error(LGBMHIP_INVALID_PARAMETER, "tree index out of range");
    const HostTree& tr = ctx->c.trees[t];
    for (int i = 0; i < tr.num_leaves - 1; ++i) {
      split_feature[i] = tr.split_feature[i];
      threshold_in_bin[i] = tr.threshold_in_bin[i];
      threshold[i] = tr.threshold[i];
      decision_type[i] = tr.decision_type[i];
      left_child[i] = tr.left_child[i];
      right_child[i] = tr.right_child[i];
      split_gain[i] = tr.split_gain[i];
      internal_value[i] = tr.internal_value[i];
    }
    for (int i = 0; i < tr.num_leaves; ++i) {
      leaf_value[i] = tr.leaf_value[i];
      leaf_count[i] = tr.leaf_count[i];
      leaf_weight[i] = tr.leaf_weight[i];
    }
  });
}
lgbmhip_status lgbmhip_leaf_of_row(lgbmhip_ctx* ctx, int32_t* out) {
  return guarded(ctx, [&] {
    Ctx* c = &ctx->c;
    if (!c->state_ready) throw Error(LGBMHIP_INVALID_PARAMETER, "no tree trained");
    set_device(c);
    DevBuf<int> d;
    d.alloc(c->N);
    // the last tree grown: class K - 1's
    WalkArgs w = walk_args(c, lane_of(c, c->K - 1), 2);
    w.out = d.p;
    launch_walk(w, grid_for(c->N, 256, c->cus * 4), c->stream);
    HIP_CHECK(hipMemcpyAsync(out, d.p, c->N * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}
lgbmhip_status lgbmhip_add_leaf_outputs(lgbmhip_ctx* ctx, const double* leaf_values, int32_t num_values,
                                        double* score) {
  return guarded(ctx, [&] {
    Ctx* c = &ctx->c;
    if (!c->state_ready || c->trees.empty()) throw Error(LGBMHIP_INVALID_PARAMETER, "no tree trained");
    if (!leaf_values || !score) throw Error(LGBMHIP_INVALID_PARAMETER, "no leaf values or scores");
    const int nl = c->trees.back().num_leaves;
    if (num_values < nl) throw Error(LGBMHIP_INVALID_PARAMETER, "fewer leaf values than the tree's leaves");
    set_device(c);
    DevBuf<int> leaf;
    DevBuf<double> lv, sc;
    leaf.alloc(c->N);
    lv.alloc(nl);
    sc.alloc(c->N);
    // each row's leaf in the last tree grown (class K - 1's), as the partition left it
    WalkArgs w = walk_args(c, lane_of(c, c->K - 1), 2);
    w.out = leaf.p;
    launch_walk(w, grid_for(c->N, 256, c->cus * 4), c->stream);
    HIP_CHECK(hipMemcpyAsync(lv.p, leaf_values, (size_t)nl * 8, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipMemcpyAsync(sc.p, score, (size_t)c->N * 8, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_add_leaf_outputs, dim3(grid_for(c->N, 256, c->cus * 4)), dim3(256), 0, c->stream, leaf.p,
                       lv.p, sc.p, (long long)c->N);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(score, sc.p, (size_t)c->N * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}
lgbmhip_status lgbmhip_copy_scores(lgbmhip_ctx* ctx, double* out) {
  return guarded(ctx, [&] {
    Ctx* c = &ctx->c;
    if (!c->state_ready) setup_training(c);
    HIP_CHECK(hipMemcpyAsync(out, c->score.p, c->score.n * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}
lgbmhip_status lgbmhip_copy_gradients(lgbmhip_ctx* ctx, float* grad, float* hess) {
  return guarded(ctx, [&] {
    Ctx* c = &ctx->c;
    if (!c->state_ready) throw Error(LGBMHIP_INVALID_PARAMETER, "no training state");
    HIP_CHECK(hipMemcpyAsync(grad, c->g.p, c->g.n * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipMemcpyAsync(hess, c->h.p, c->h.n * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}
lgbmhip_status lgbmhip_quant_state(lgbmhip_ctx* ctx, int8_t* gq, int8_t* hq, double* scales) {
  return guarded(ctx, [&] {
    Ctx* c = &ctx->c;
    if (!c->state_ready || !c->opts.use_quantized_grad) throw Error(LGBMHIP_INVALID_PARAMETER, "no quantized state");
    // buffer 0 of the last tree's lane holds its (gq, hq) pairs in row order
    std::vector<int8_t> q((size_t)2 * c->N);
    HIP_CHECK(hipMemcpyAsync(q.data(), lane_of(c, c->K - 1).gh[0].p, q.size(), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < c->N; ++i) {
      gq[i] = q[2 * i];
      hq[i] = q[2 * i + 1];
    }
    IterDev it;
    HIP_CHECK(hipMemcpyAsync(&it, c->it.p, sizeof(IterDev), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    scales[0] = it.grad_scale;
    scales[1] = it.hess_scale;
  });
}

lgbmhip_status lgbmhip_gradients(lgbmhip_ctx* ctx, const double* scores, float* grad, float* hess) {
  return guarded(ctx, [&] {
    Ctx* c = &ctx->c;
    if (!c->state_ready) setup_training(c);
    set_device(c);
    HIP_CHECK(hipMemcpyAsync(c->score.p, scores, c->score.n * 8, hipMemcpyHostToDevice, c->stream));
    enqueue_gradients(c);
    HIP_CHECK(hipMemcpyAsync(grad, c->g.p, c->g.n * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipMemcpyAsync(hess, c->h.p, c->h.n * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

lgbmhip_status lgbmhip_discretize(lgbmhip_ctx* ctx, const float* grad, const float* hess, int32_t num_calls,
                                  int8_t* gq, int8_t* hq, double* scales) {
  return guarded(ctx, [&] {
    Ctx* c = &ctx->c;
    if (!c->opts.use_quantized_grad) throw Error(LGBMHIP_INVALID_PARAMETER, "use_quantized_grad is off");
    if (!c->state_ready) setup_training(c);
    set_device(c);
    const int64_t N = c->N;
    if (num_calls < 1 || c->n_calls + num_calls > (int)c->starts.n)
      throw Error(LGBMHIP_INVALID_PARAMETER, "num_calls out of range of the start-offset table");
    HIP_CHECK(hipMemcpyAsync(c->g.p, grad, N * 4, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipMemcpyAsync(c->h.p, hess, N * 4, hipMemcpyHostToDevice, c->stream));
    for (int call = 0; call < num_calls; ++call) {
      hipLaunchKernelGGL(k_iter_reset, dim3(1), dim3(64), 0, c->stream, c->it.p, 1);
      hipLaunchKernelGGL(k_absmax_it, dim3(grid_for(N, 256, c->cus * 4)), dim3(256), 0, c->stream, c->g.p, c->h.p,
                         (long long)N, c->it.p, (const unsigned*)nullptr);
      hipLaunchKernelGGL(k_discretize, dim3(grid_for(N, 256, c->cus * 4)), dim3(256), 0, c->stream, (long long)N,
                         c->g.p, c->h.p, c->grv.p, c->hrv.p, c->opts.num_grad_quant_bins, c->opts.stochastic_rounding,
                         0, c->starts.p, c->call_ctr.p, 0, c->it.p, (unsigned*)nullptr, (char2*)c->lanes[0]->gh[0].p,
                         c->gq.p, c->hq.p, (const unsigned*)nullptr);
      hipLaunchKernelGGL(k_bump, dim3(1), dim3(1), 0, c->stream, c->call_ctr.p, 1);
      c->n_calls++;
    }
    HIP_CHECK(hipMemcpyAsync(gq, c->gq.p, N, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipMemcpyAsync(hq, c->hq.p, N, hipMemcpyDeviceToHost, c->stream));
    IterDev it;
    HIP_CHECK(hipMemcpyAsync(&it, c->it.p, sizeof(IterDev), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    scales[0] = it.grad_scale;
    scales[1] = it.hess_scale;
  });
}

static void stage_hist(Ctx* c, const uint32_t* rows, int64_t m, const void* gsrc, const void* hsrc, bool quant,
                       void* out, int32_t* out_B) {
  if (!c->state_ready) setup_training(c);
  set_device(c);
  if (quant != (c->opts.use_quantized_grad != 0)) throw Error(LGBMHIP_INVALID_PARAMETER, "histogram kind differs from train opts");
  Lane& ln = *c->lanes[0];   // lane 0 runs on the context's stream
  const int64_t N = c->N;
  if (m < 0 || m > N) throw Error(LGBMHIP_INVALID_DATA, "row count out of range");
  for (int64_t i = 0; i < m; ++i)
    if (rows[i] >= (uint64_t)N) throw Error(LGBMHIP_INVALID_DATA, "row index out of range");
  // pack (g, h) in the row order of `rows` into gh buffer 1 positions 0..m-1
  std::vector<uint8_t> pk((size_t)std::max<int64_t>(m, 1) * (quant ? 2 : 8));
  unsigned gbits = 0, hbits = 0;
  int qg = 0, qh = 0;   // max |gq| and max u8(hq) of the rows
  for (int64_t i = 0; i < m; ++i) {
    const uint32_t r = rows[i];
    if (quant) {
      pk[2 * i] = ((const uint8_t*)gsrc)[r];
      pk[2 * i + 1] = ((const uint8_t*)hsrc)[r];
      qg = std::max(qg, std::abs((int)((const int8_t*)gsrc)[r]));
      qh = std::max(qh, (int)((const uint8_t*)hsrc)[r]);
    } else {
      float gv = ((const float*)gsrc)[r], hv = ((const float*)hsrc)[r];
      std::memcpy(&pk[8 * i], &gv, 4);
      std::memcpy(&pk[8 * i + 4], &hv, 4);
      unsigned a, b;
      float ga = std::fabs(gv), ha = std::fabs(hv);
      std::memcpy(&a, &ga, 4);
      std::memcpy(&b, &ha, 4);
      gbits = std::max(gbits, a);
      hbits = std::max(hbits, b);
    }
  }
  if (m > 0) {
    // the rows' bins gathered into positions 0..m-1 of ordered buffer 1
    c->rows_buf.alloc(m);
    HIP_CHECK(hipMemcpyAsync(c->rows_buf.p, rows, m * 4, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipMemcpyAsync(ln.gh[1].p, pk.data(), pk.size(), hipMemcpyHostToDevice, c->stream));
    launch_group_rows(c, c->rows_buf.p, m, ln.grp[1].p, ln.grp[1].n, c->stream);
  }
  HistArgs a = hist_args(c, ln, quant, 0, 1);
  a.d_begin = 0;
  a.d_count = m;
  a.d_buf = 1;
  a.d_shift_g = fix_shift(m, gbits);
  a.d_shift_h = fix_shift(m, hbits);
  a.q_host = 1;
  a.q_gmax = qg;
  a.q_hmax = qh;
  const int C = quant ? 1 : 2;
  const size_t cnt = (size_t)c->F * c->B * C;
  std::vector<long long> acc(ln.acc.n, 0);
  if (m > 0) {
    launch_hist(a, quant, c->hist_tpr, c->hist_chunks * c->hist_ng, c->hist_lds, c->stream);
    slab_total(c, ln, quant, c->stream);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(acc.data(), ln.acc.p, ln.acc.n * 8, hipMemcpyDeviceToHost, c->stream));
  }
  HIP_CHECK(hipStreamSynchronize(c->stream));
  // acc is [group][feature in group][bin][C]: out is [feature][bin][C]
  for (size_t i = 0; i < cnt; ++i) {
    const size_t f = i / ((size_t)c->B * C), r = i % ((size_t)c->B * C);
    const FeatDev& fd = c->feats[f];
    const bool in = r / C < (size_t)fd.num_bin;   // a 4-bit feature's unit holds its 16 bins only
    const long long v = in ? acc[(size_t)fd.hg * c->hist_pu * C + (size_t)fd.hoff * C + r] : 0;
    if (quant) ((long long*)out)[i] = v;
    else ((double*)out)[i] = std::ldexp((double)v, -((i & 1) ? a.d_shift_h : a.d_shift_g));
  }
  *out_B = c->B;
}

lgbmhip_status lgbmhip_build_histogram(lgbmhip_ctx* ctx, const uint32_t* rows, int64_t m, const float* grad,
                                       const float* hess, double* out, int32_t* out_num_bins) {
  return guarded(ctx, [&] { stage_hist(&ctx->c, rows, m, grad, hess, false, out, out_num_bins); });
}
lgbmhip_status lgbmhip_build_histogram_int(lgbmhip_ctx* ctx, const uint32_t* rows, int64_t m, const int8_t* gq,
                                           const int8_t* hq, int64_t* out, int32_t* out_num_bins) {
  return guarded(ctx, [&] { stage_hist(&ctx->c, rows, m, gq, hq, true, out, out_num_bins); });
}

static void stage_search(Ctx* c, int f, const double* hist, const int64_t* ihist, LeafSums ls, bool quant,
                         double* od, int32_t* oi, int64_t* og, uint32_t* ocat) {
  if (f < 0 || f >= c->F) throw Error(LGBMHIP_INVALID_PARAMETER, "feature out of range");
  if (!c->sp_set) throw Error(LGBMHIP_INVALID_PARAMETER, "split params not set");
  set_device(c);
  const int nb = c->feats[f].num_bin;
  DevBuf<double> dh;
  DevBuf<long long> dih;
  DevBuf<SplitRec> dr;
  DevBuf<int> ds;
  dr.alloc(1);
  ds.alloc(1);
  if (quant) {
    dih.alloc(nb);
    HIP_CHECK(hipMemcpy(dih.p, ihist, nb * 8, hipMemcpyHostToDevice));
  } else {
    dh.alloc(2 * nb);
    HIP_CHECK(hipMemcpy(dh.p, hist, nb * 16, hipMemcpyHostToDevice));
  }
  hipLaunchKernelGGL(k_search_one, dim3(1), dim3(64), 0, c->stream, dh.p, dih.p, c->feats[f], c->sp, ls, quant ? 1 : 0,
                     dr.p, ds.p);
  SplitRec r;
  int spl = 0;
  HIP_CHECK(hipMemcpyAsync(&r, dr.p, sizeof(r), hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipMemcpyAsync(&spl, ds.p, 4, hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  od[0] = r.gain; od[1] = r.left_output; od[2] = r.right_output;
  od[3] = r.lsg; od[4] = r.lsh; od[5] = r.rsg; od[6] = r.rsh;
  oi[0] = r.threshold; oi[1] = r.default_left; oi[2] = r.left_count; oi[3] = r.right_count; oi[4] = spl;
  if (og) { og[0] = r.lgh; og[1] = r.rgh; }
  if (ocat)
    for (int w = 0; w < kCatWords; ++w) ocat[w] = spl && c->feats[f].bin_type == 1 ? r.cat_bits[w] : 0u;
}

lgbmhip_status lgbmhip_find_best_split(lgbmhip_ctx* ctx, int32_t feature, const double* hist, double sum_g,
                                       double sum_h, int32_t num_data, double parent_output, double* out_d,
                                       int32_t* out_i, uint32_t* out_cat_bits) {
  return guarded(ctx, [&] {
    LeafSums ls{};
    ls.sg = sum_g;
    ls.sh = sum_h + 2.0 * (double)1e-15f;
    ls.nd = num_data;
    ls.po = parent_output;
    stage_search(&ctx->c, feature, hist, nullptr, ls, false, out_d, out_i, nullptr, out_cat_bits);
  });
}
lgbmhip_status lgbmhip_find_best_split_int(lgbmhip_ctx* ctx, int32_t feature, const int64_t* hist, int64_t int_sum,
                                           double grad_scale, double hess_scale, int32_t num_data,
                                           double parent_output, double* out_d, int32_t* out_i, int64_t* out_gh,
                                           uint32_t* out_cat_bits) {
  return guarded(ctx, [&] {
    LeafSums ls{};
    ls.isum = int_sum;
    ls.gs = grad_scale;
    ls.hs = hess_scale;
    ls.nd = num_data;
    ls.po = parent_output;
    stage_search(&ctx->c, feature, nullptr, hist, ls, true, out_d, out_i, out_gh, out_cat_bits);
  });
}

// The stable partition of device rows d_rows[0..m) into d_out (left rows in
// order, then right rows in order) with the kSpartTile kernels on the
// context's stream; work: at least ntiles + 2 words of device memory. Returns
// the left count; a row index >= N is an INVALID_DATA error (the kernels
// read nothing for it).
static int64_t stable_partition(Ctx* c, const unsigned* d_rows, int64_t m, int feature, uint32_t threshold,
                                int default_left, unsigned* d_out, unsigned* work) {
  const int ntiles = (int)((m + kSpartTile - 1) / kSpartTile);
  SplitRec s{};
  s.threshold = (int)threshold;
  s.default_left = default_left;
  const uint8_t* col = c->bins_cm.p + (size_t)feature * c->N;
  // the out-of-range flag word starts at zero on every call (a fill kernel, no memset node)
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, c->stream, work + ntiles + 1, 1LL, 0u);
  if (m > 0) {
    hipLaunchKernelGGL(k_spart_count, dim3(ntiles), dim3(256), 0, c->stream, d_rows, (long long)m, (unsigned)c->N,
                       col, c->feats[feature], s, work, ntiles);
    hipLaunchKernelGGL(k_spart_scan, dim3(1), dim3(1024), 0, c->stream, work, ntiles);
    hipLaunchKernelGGL(k_spart_scatter, dim3(ntiles), dim3(256), 0, c->stream, d_rows, (long long)m,
                       (unsigned)c->N, col, c->feats[feature], s, work, ntiles, d_out);
  } else {
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, c->stream, work, 1LL, 0u);
  }
  HIP_CHECK(hipGetLastError());
  unsigned tail[2] = {0, 0};   // {left rows, out-of-range flag}
  HIP_CHECK(hipMemcpyAsync(tail, work + ntiles, sizeof tail, hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  if (tail[1]) throw Error(LGBMHIP_INVALID_DATA, "row index out of range");
  return (int64_t)tail[0];
}

static void check_partition_args(Ctx* c, int64_t m, int32_t feature) {
  if (!c->bins_cm.p) throw Error(LGBMHIP_INVALID_PARAMETER, "no bins uploaded");
  if (feature < 0 || feature >= c->F) throw Error(LGBMHIP_INVALID_PARAMETER, "feature out of range");
  if (m < 0 || m > c->N) throw Error(LGBMHIP_INVALID_DATA, "row count out of range");
  if (c->feats[feature].bin_type != 0) throw Error(LGBMHIP_UNSUPPORTED, "partition on a categorical feature");
}

lgbmhip_status lgbmhip_partition(lgbmhip_ctx* ctx, const uint32_t* rows, int64_t m, int32_t feature,
                                 uint32_t threshold, int32_t default_left, uint32_t* out_rows, int64_t* out_left) {
  return guarded(ctx, [&] {
    Ctx* c = &ctx->c;
    check_partition_args(c, m, feature);
    set_device(c);
    for (int64_t i = 0; i < m; ++i)
      if (rows[i] >= (uint64_t)c->N) throw Error(LGBMHIP_INVALID_DATA, "row index out of range");
    DevBuf<unsigned> dr, dout, work;
    dr.alloc(std::max<int64_t>(m, 1));
    dout.alloc(std::max<int64_t>(m, 1));
    work.alloc((size_t)(m + kSpartTile - 1) / kSpartTile + 2);
    if (m > 0) HIP_CHECK(hipMemcpyAsync(dr.p, rows, m * 4, hipMemcpyHostToDevice, c->stream));
    *out_left = stable_partition(c, dr.p, m, feature, threshold, default_left, dout.p, work.p);
    if (m > 0) HIP_CHECK(hipMemcpyAsync(out_rows, dout.p, m * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

lgbmhip_status lgbmhip_partition_device_noalloc(lgbmhip_ctx* ctx, const uint32_t* d_rows, int64_t m, int32_t feature,
                                                uint32_t threshold, int32_t default_left, uint32_t* d_out_rows,
                                                int64_t* out_left) {
  return guarded(ctx, [&] {
    Ctx* c = &ctx->c;
    if (m > kNoAllocMaxRows)
      throw Error(LGBMHIP_INVALID_PARAMETER, "lgbmhip_partition_device_noalloc takes at most " +
                                                 std::to_string(kNoAllocMaxRows) + " rows");
    check_partition_args(c, m, feature);
    if (m > 0 && (!d_rows || !d_out_rows)) throw Error(LGBMHIP_INVALID_PARAMETER, "no device rows");
    if (!out_left) throw Error(LGBMHIP_INVALID_PARAMETER, "no output for the left count");
    set_device(c);
    unsigned* work = scratch_of(c->device);
    *out_left = stable_partition(c, d_rows, m, feature, threshold, default_left, d_out_rows, work);
  });
}

lgbmhip_status lgbmhip_predict_raw(lgbmhip_ctx* ctx, const float* X, int64_t rows, int32_t ncol, double* out) {
  return guarded(ctx, [&] { predict(&ctx->c, X, nullptr, rows, ncol, out); });
}
lgbmhip_status lgbmhip_predict_binned(lgbmhip_ctx* ctx, const uint8_t* bins, int64_t rows, double* out) {
  return guarded(ctx, [&] { predict(&ctx->c, nullptr, bins, rows, 0, out); });
}
lgbmhip_status lgbmhip_predict_device(lgbmhip_ctx* ctx, const float* d_X, const uint8_t* d_bins, int64_t rows,
                                      int32_t ncol, double* d_out) {
  return guarded(ctx, [&] {
    Ctx* c = &ctx->c;
    if ((d_X == nullptr) == (d_bins == nullptr)) throw Error(LGBMHIP_INVALID_PARAMETER, "exactly one of X, bins");
    if (!d_out || rows < 0 || (d_X && ncol <= 0)) throw Error(LGBMHIP_INVALID_PARAMETER, "bad prediction buffers");
    set_device(c);
    launch_predict(c, d_X, d_bins, rows, ncol, d_out);
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}
lgbmhip_status lgbmhip_upload_raw(lgbmhip_ctx* ctx, const lgbmhip_bins* mappers, const float* X, int64_t rows,
                                  int32_t ncol, int32_t x_on_device) {
  return guarded(ctx, [&] {
    if (!mappers || !X) throw Error(LGBMHIP_INVALID_PARAMETER, "mappers and rows are required");
    upload_raw(&ctx->c, *mappers->b, X, rows, ncol, x_on_device != 0);
    ctx->c.num_groups = mappers->b->num_groups;
  });
}
lgbmhip_status lgbmhip_bin_rows_device(lgbmhip_ctx* ctx, const float* d_X, int64_t rows, int32_t ncol,
                                       uint8_t* d_out) {
  return guarded(ctx, [&] {
    Ctx* c = &ctx->c;
    if (!c->N) throw Error(LGBMHIP_INVALID_DATA, "upload the training data first");
    if (!d_X || !d_out || rows < 0) throw Error(LGBMHIP_INVALID_PARAMETER, "bad device buffers");
    set_device(c);
    launch_bin_rows(c, d_X, rows, ncol, d_out, c->F, false);
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}
lgbmhip_status lgbmhip_add_valid(lgbmhip_ctx* ctx, const uint8_t* bins, int64_t rows, const float* label,
                                 const float* weight, const int32_t* query_boundaries, int32_t num_queries,
                                 const double* init_score, int32_t* out_index) {
  return guarded(ctx, [&] {
    const int idx = add_valid(&ctx->c, bins, rows, label, weight, query_boundaries, num_queries, init_score);
    if (out_index) *out_index = idx;
  });
}
lgbmhip_status lgbmhip_copy_valid_scores(lgbmhip_ctx* ctx, int32_t index, double* out) {
  return guarded(ctx, [&] {
    Ctx* c = &ctx->c;
    if (index < 1 || index > (int32_t)c->valid.size()) throw Error(LGBMHIP_INVALID_PARAMETER, "no such validation set");
    set_device(c);
    const ValidData& v = *c->valid[index - 1];
    HIP_CHECK(hipMemcpyAsync(out, v.score.p, v.score.n * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}
lgbmhip_status lgbmhip_eval(lgbmhip_ctx* ctx, int32_t data, const lgbmhip_metric_opts* opts, double* out,
                            int32_t* out_count) {
  return guarded(ctx, [&] {
    if (!opts || !out) throw Error(LGBMHIP_INVALID_PARAMETER, "metric options and output are required");
    const int cnt = eval_metric(&ctx->c, data, opts, out);
    if (out_count) *out_count = cnt;
  });
}
// A string result: its length always, the text (NUL-terminated) when it fits.
static void put_text(const std::string& s, char* out, int64_t capacity, int64_t* out_len) {
  if (out_len) *out_len = (int64_t)s.size();
  if (!out) return;
  if (capacity <= (int64_t)s.size()) throw Error(LGBMHIP_INVALID_PARAMETER, "output buffer too small for the text");
  std::memcpy(out, s.c_str(), s.size() + 1);
}
lgbmhip_status lgbmhip_predict_leaf(lgbmhip_ctx* ctx, const float* X, int64_t rows, int32_t ncol, double* out,
                                    int32_t* out_trees) {
  return guarded(ctx, [&] {
    if (!X || !out || rows < 0 || ncol <= 0) throw Error(LGBMHIP_INVALID_PARAMETER, "bad prediction buffers");
    const int64_t T = predict_leaf(&ctx->c, X, rows, ncol, out);
    if (out_trees) *out_trees = (int32_t)T;
  });
}
lgbmhip_status lgbmhip_rollback_one_iter(lgbmhip_ctx* ctx) {
  return guarded(ctx, [&] { rollback_one_iter(&ctx->c); });
}
lgbmhip_status lgbmhip_set_predict_window(lgbmhip_ctx* ctx, int32_t start_iteration, int32_t num_iteration,
                                          int32_t early_stop_freq, double early_stop_margin) {
  return guarded(ctx, [&] {
    // Gbdt::prediction_early_stop (predict.rs:178-190)
    if (early_stop_freq < 0) throw Error(LGBMHIP_INVALID_PARAMETER, "Check failed: (early_stop_freq) > (0)");
    if (early_stop_freq > 0 && !(early_stop_margin >= 0.0))
      throw Error(LGBMHIP_INVALID_PARAMETER, "Check failed: (early_stop_margin) >= (0)");
    Ctx* c = &ctx->c;
    c->pred_start = start_iteration;
    c->pred_num = num_iteration;
    c->pred_es_freq = early_stop_freq;
    c->pred_es_margin = early_stop_margin;
  });
}
lgbmhip_status lgbmhip_save_model(lgbmhip_ctx* ctx, int32_t num_iteration, const char* feature_names,
                                  const char* feature_infos, const char* parameters, char* out, int64_t capacity,
                                  int64_t* out_len) {
  return guarded(ctx, [&] {
    put_text(save_model(&ctx->c, num_iteration, feature_names, feature_infos, parameters), out, capacity, out_len);
  });
}
lgbmhip_status lgbmhip_check_model(const char* text, int32_t* out_num_trees, int32_t* out_num_tree_per_iteration,
                                   char* out, int64_t capacity, int64_t* out_len) {
  return guarded(nullptr, [&] {
    const ParsedModel m = parse_model(text);
    if (out_num_trees) *out_num_trees = (int32_t)m.trees.size();
    if (out_num_tree_per_iteration) *out_num_tree_per_iteration = m.head.ntpi;
    if (out || out_len) put_text(model_text(m.head, m.trees, 0), out, capacity, out_len);
  });
}
lgbmhip_status lgbmhip_load_model(lgbmhip_ctx* ctx, const char* text) {
  return guarded(ctx, [&] { load_model(&ctx->c, text); });
}
lgbmhip_status lgbmhip_bins_feature_infos(const lgbmhip_bins* b, char* out, int64_t capacity, int64_t* out_len) {
  return guarded(nullptr, [&] {
    // Dataset::feature_infos (dataset.rs:770-776), BinMapper::bin_info_string (binning.rs:680-686)
    std::string s;
    for (int c = 0; c < b->b->ncol; ++c) {
      const Mapper& m = b->b->mappers[c];
      if (c) s += ' ';
      if (m.is_trivial) {
        s += "none";
      } else if (m.bin_type == 1) {
        for (size_t i = 0; i < m.categories.size(); ++i) s += (i ? ":" : "") + std::to_string(m.categories[i]);
      } else {
        s += "[" + fmt_g(m.min_val, 17) + ":" + fmt_g(m.max_val, 17) + "]";
      }
    }
    put_text(s, out, capacity, out_len);
  });
}
lgbmhip_status lgbmhip_profile(lgbmhip_ctx* ctx, int32_t enable, double* out12) {
  return guarded(ctx, [&] {
    Ctx* c = &ctx->c;
    if (out12) for (int i = 0; i < 12; ++i) out12[i] = c->prof[i];
    if (enable >= 0) {
      c->profile = enable;
      for (int i = 0; i < 12; ++i) c->prof[i] = 0.0;
      if (enable && c->pev.empty()) {
        set_device(c);
        if (!c->state_ready) setup_training(c);
        c->pev.resize((size_t)c->K * prof_steps(c) * 4 * 2);
        for (auto& e : c->pev) HIP_CHECK(hipEventCreate(&e));
      }
    }
  });
}

lgbmhip_status lgbmhip_lds_atomic_peak(lgbmhip_ctx* ctx, double* lane_atomics_per_sec) {
  return guarded(ctx, [&] {
    Ctx* c = &ctx->c;
    if (!lane_atomics_per_sec) throw Error(LGBMHIP_INVALID_PARAMETER, "null argument");
    set_device(c);
    const int blocks = 3 * c->cus, iters = 4096;
    DevBuf<unsigned> sink;
    sink.alloc((size_t)blocks * 64);
    struct Events {   // destroyed on every exit
      hipEvent_t e[2] = {nullptr, nullptr};
      ~Events() {
        for (hipEvent_t x : e)
          if (x) (void)hipEventDestroy(x);
      }
    } ev;
    HIP_CHECK(hipEventCreate(&ev.e[0]));
    HIP_CHECK(hipEventCreate(&ev.e[1]));
    hipLaunchKernelGGL(k_lds_atomic_peak, dim3(blocks), dim3(512), 0, c->stream, 64, sink.p);   // warm
    HIP_CHECK(hipEventRecord(ev.e[0], c->stream));
    hipLaunchKernelGGL(k_lds_atomic_peak, dim3(blocks), dim3(512), 0, c->stream, iters, sink.p);
    HIP_CHECK(hipEventRecord(ev.e[1], c->stream));
    HIP_CHECK(hipEventSynchronize(ev.e[1]));
    HIP_CHECK(hipGetLastError());
    float ms = 0.0f;
    HIP_CHECK(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]));
    if (!(ms > 0.0f)) throw Error(LGBMHIP_INTERNAL, "LDS-atomic timing failed");
    *lane_atomics_per_sec = (double)blocks * 512.0 * iters * 16.0 / (ms * 1e-3);
  });
}

lgbmhip_status lgbmhip_set_tuning(lgbmhip_ctx* ctx, int32_t part_grid, int32_t hist_chunks, int32_t hist_min_rows,
                                  int32_t hist_lds_budget, int32_t rounds_main, int32_t rounds_cont,
                                  int32_t late_round) {
  return guarded(ctx, [&] {
    Ctx* c = &ctx->c;
    if (part_grid < 0 || hist_chunks < 0 || hist_min_rows < 0 || hist_lds_budget < 0 || rounds_main < 0 ||
        rounds_cont < 0 || late_round < -1)
      throw Error(LGBMHIP_INVALID_PARAMETER, "tuning values must be >= 0 (late_round >= -1)");
    if (late_round >= 0) c->tune_late_round = late_round;
    if (hist_lds_budget > 0 && hist_lds_budget < 4096) throw Error(LGBMHIP_INVALID_PARAMETER, "LDS budget below 4 KiB");
    if (rounds_main > 4096 || rounds_cont > 4096) throw Error(LGBMHIP_INVALID_PARAMETER, "growth rounds above 4096");
    if (part_grid) c->tune_part_grid = part_grid;
    if (hist_chunks) c->tune_hist_chunks = hist_chunks;
    if (hist_min_rows) c->tune_hist_min_rows = hist_min_rows;
    if (hist_lds_budget) c->tune_hist_lds = hist_lds_budget;
    if (rounds_main) c->rounds_main = rounds_main;
    if (rounds_cont) c->rounds_cont = rounds_cont;
    c->state_ready = false;   // shapes are baked into the captured graphs
  });
}

lgbmhip_status lgbmhip_autotune(lgbmhip_ctx* ctx, int32_t trees_per_candidate, int32_t* out5) {
  return guarded(ctx, [&] {
    const auto r = autotune(&ctx->c, trees_per_candidate);
    if (out5)
      for (int i = 0; i < 5; ++i) out5[i] = r[(size_t)i];
  });
}
lgbmhip_status lgbmhip_set_feature_groups(lgbmhip_ctx* ctx, int32_t num_groups) {
  return guarded(ctx, [&] {
    if (num_groups < 0) throw Error(LGBMHIP_INVALID_PARAMETER, "the feature-group count must be non-negative");
    ctx->c.num_groups = num_groups;
  });
}
lgbmhip_status lgbmhip_set_bin_packing(lgbmhip_ctx* ctx, int32_t u4) {
  return guarded(ctx, [&] {
    if (u4 != 0 && u4 != 1) throw Error(LGBMHIP_INVALID_PARAMETER, "bin packing must be 0 or 1");
    ctx->c.pack_u4 = u4;
  });
}
lgbmhip_status lgbmhip_bin_layout(lgbmhip_ctx* ctx, int32_t* out3) {
  return guarded(ctx, [&] {
    const Ctx* c = &ctx->c;
    int n4 = 0;
    for (const FeatDev& d : c->feats) n4 += d.nib >= 0 ? 1 : 0;
    out3[0] = c->W;
    out3[1] = c->S;
    out3[2] = n4;
  });
}
lgbmhip_status lgbmhip_get_tuning(lgbmhip_ctx* ctx, int32_t* out10) {
  return guarded(ctx, [&] {
    Ctx* c = &ctx->c;
    if (!c->state_ready) setup_training(c);
    out10[0] = c->part_grid;
    out10[1] = c->hist_chunks;
    out10[2] = c->hist_min_rows;
    out10[3] = (int32_t)c->hist_lds;
    out10[4] = c->hist_fb;
    out10[5] = c->hist_ng;
    out10[6] = c->hist_tpr;
    out10[7] = c->rounds_main;
    out10[8] = c->rounds_cont;
    out10[9] = c->late_round;
  });
}

lgbmhip_status lgbmhip_last_timings(lgbmhip_ctx* ctx, double* out4) {
  return guarded(ctx, [&] {
    out4[0] = ctx->c.last_ms;
    out4[1] = out4[2] = out4[3] = 0.0;
  });
}

}  // extern "C"
