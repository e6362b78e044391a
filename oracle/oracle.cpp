// ORACLE — test infrastructure only.
//
// A single-threaded CPU restatement of the reference engine's GBDT hot path
// (zrluety/lightgbm-rust, crate lgbm-core), used as the correctness checker
// for the HIP engine. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load this library; the product never links or calls it.
//
// Every function cites the reference file:line it follows (paths relative
// to crates/lgbm-core/src/). Semantics are those of one thread
// (`num_threads = 1`, `nblock == 1`): sequential sums in row order.
//
// Parity status of each part (what pins it):
//   Random / Mt19937           : known answers (random.rs:97-105, mt19937.rs:95-103)
//   BinMapper                  : binning.rs:704-779 known answers + brute-force bin invariants
//   histogram fix/subtract     : histogram.rs:91-107 known answers + totals invariant
//   find_best_threshold        : split.rs:1423-1493 known answers + brute force over thresholds
//   quantizer                  : quant.rs:166-184 known answers
//   partition                  : std::stable_partition equivalence (partition.rs:1-7)
//   learner                    : learner/tests.rs:24-89 known answers
//   binary/multiclass gradients: closed forms (binary.rs:101-127, multiclass.rs:126-154)
//   lambdarank gradients       : two-document known answer (objective/rank.rs:491-509)
//   AUC                        : sklearn roc_auc_score on tie-free and tied inputs
//   metric_auc (weighted)      : sklearn roc_auc_score(sample_weight) + brute force over pairs
//   metric_point               : sklearn log_loss / accuracy (binary and multiclass, weighted)
//   metric_ndcg                : dcg.rs:195-209 known answers + sklearn ndcg_score (gains 2^rel-1)
//   feature-group count        : cases the greedy bundling fixes (dense, exclusive one-hot, sparse merge)
//   full-pipeline model text   : parity unpinned (no committed artifact in the reference)

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <omp.h>
#include <set>
#include <string>
#include <unordered_map>
#include <vector>

namespace orc {

// consts.rs:8-16
static const double K_EPSILON = (double)1e-15f;
static const double K_ZERO_THRESHOLD = (double)1e-35f;
static const double K_MIN_SCORE = -INFINITY;
static const double K_SPARSE_THRESHOLD = 0.7;

// ---------------------------------------------------------------- RNGs
// random.rs:10-71 (upstream LCG)
struct Random {
  uint32_t x;
  explicit Random(int seed) : x((uint32_t)seed) {}
  int rand_int16() { x = x * 214013u + 2531011u; return (int)((x >> 16) & 0x7FFF); }
  int rand_int32() { x = x * 214013u + 2531011u; return (int)(x & 0x7FFFFFFF); }
  int next_short(int lo, int hi) { return rand_int16() % (hi - lo) + lo; }
  int next_int(int lo, int hi) { return rand_int32() % (hi - lo) + lo; }
  float next_float() { return (float)rand_int16() / 32768.0f; }
  std::vector<int> sample(int n, int k) {
    std::vector<int> ret;
    if (k > n || k <= 0) return ret;
    if (k == n) {
      for (int i = 0; i < n; ++i) ret.push_back(i);
    } else if (k > 1 && (double)k > ((double)n / std::log2((double)k))) {
      for (int i = 0; i < n; ++i) {
        double prob = (double)(k - (int)ret.size()) / (double)(n - i);
        if ((double)next_float() < prob) ret.push_back(i);
      }
    } else {
      std::set<int> s;
      for (int r = n - k; r < n; ++r) {
        int v = next_int(0, r + 1);
        if (!s.insert(v).second) s.insert(r);
      }
      ret.assign(s.begin(), s.end());
    }
    return ret;
  }
};

// mt19937.rs:14-88
struct Mt19937 {
  uint32_t st[624];
  int idx;
  explicit Mt19937(uint32_t seed) {
    st[0] = seed;
    for (int i = 1; i < 624; ++i) st[i] = 1812433253u * (st[i - 1] ^ (st[i - 1] >> 30)) + (uint32_t)i;
    idx = 624;
  }
  void twist() {
    for (int k = 0; k < 624; ++k) {
      uint32_t y = (st[k] & 0x80000000u) | (st[(k + 1) % 624] & 0x7fffffffu);
      uint32_t mag = (y & 1u) ? 0x9908b0dfu : 0u;
      st[k] = st[(k + 397) % 624] ^ (y >> 1) ^ mag;
    }
    idx = 0;
  }
  uint32_t next_u32() {
    if (idx >= 624) twist();
    uint32_t z = st[idx++];
    z ^= z >> 11;
    z ^= (z << 7) & 0x9d2c5680u;
    z ^= (z << 15) & 0xefc60000u;
    z ^= z >> 18;
    return z;
  }
  double canonical() {
    const double R = 4294967296.0;
    double sum = (double)next_u32();
    sum += (double)next_u32() * R;
    double ret = sum / (R * R);
    if (ret >= 1.0) { uint64_t b = 0x3FEFFFFFFFFFFFFFull; std::memcpy(&ret, &b, 8); }
    return ret;
  }
  double uniform01() { return canonical() * (1.0 - 0.0) + 0.0; }
  int uniform_int(int a, int b) {
    uint32_t range = (uint32_t)((int64_t)b - (int64_t)a + 1);
    uint64_t product = (uint64_t)next_u32() * (uint64_t)range;
    uint32_t low = (uint32_t)product;
    if (low < range) {
      uint32_t threshold = (uint32_t)(-range) % range;
      while (low < threshold) {
        product = (uint64_t)next_u32() * (uint64_t)range;
        low = (uint32_t)product;
      }
    }
    return (int)((int64_t)(product >> 32) + (int64_t)a);
  }
};

// ---------------------------------------------------------------- binning
enum MissingType { MT_NONE = 0, MT_ZERO = 1, MT_NAN = 2 };
enum BinType { BT_NUM = 0, BT_CAT = 1 };

struct BinParams {
  int max_bin, min_data_in_bin, min_split_data;
  bool pre_filter, use_missing, zero_as_missing;
};

static inline double next_up(double a) { return std::nextafter(a, INFINITY); }
static inline bool check_double_equal_ordered(double a, double b) { return b <= next_up(a); }

// binning.rs:43-64
struct BinMapper {
  int num_bin = 1;
  int missing_type = MT_NONE;
  std::vector<double> bin_upper_bound{INFINITY};
  bool is_trivial = true;
  double sparse_rate = 1.0;
  int bin_type = BT_NUM;
  double min_val = 0, max_val = 0;
  uint32_t default_bin = 0, most_freq_bin = 0;
  std::vector<int> bin_2_categorical;
  std::unordered_map<int, uint32_t> categorical_2_bin;

  // binning.rs:637-667
  uint32_t value_to_bin(double value) const {
    if (bin_type == BT_CAT) {
      if (std::isnan(value)) return 0;
      int v = (int)value;
      if (v < 0) return 0;
      auto it = categorical_2_bin.find(v);
      return it == categorical_2_bin.end() ? 0 : it->second;
    }
    if (std::isnan(value)) {
      if (missing_type == MT_NAN) return (uint32_t)(num_bin - 1);
      value = 0.0;
    }
    int l = 0, r = num_bin - 1;
    if (missing_type == MT_NAN) r -= 1;
    while (l < r) {
      int m = (r + l - 1) / 2;
      if (value <= bin_upper_bound[m]) r = m; else l = m + 1;
    }
    return (uint32_t)l;
  }
  double bin_to_value(uint32_t b) const {
    return bin_type == BT_CAT ? (double)bin_2_categorical[b] : bin_upper_bound[b];
  }
};

// binning.rs:94-114
static bool need_filter(const std::vector<int>& cnt, int total, int filter_cnt, int bin_type) {
  if (bin_type == BT_NUM) {
    int sum_left = 0;
    for (size_t i = 0; i + 1 < cnt.size(); ++i) {
      sum_left += cnt[i];
      if (sum_left >= filter_cnt && total - sum_left >= filter_cnt) return false;
    }
  } else {
    if (cnt.size() > 2) return false;
    for (size_t i = 0; i + 1 < cnt.size(); ++i) {
      int sum_left = cnt[i];
      if (sum_left >= filter_cnt && total - sum_left >= filter_cnt) return false;
    }
  }
  return true;
}

// binning.rs:117-212
static std::vector<double> greedy_find_bin(const double* dv, const int* counts, int nd, int max_bin,
                                           size_t total_cnt, int min_data_in_bin) {
  std::vector<double> ub;
  if (nd == 0) return {INFINITY};
  if (nd <= max_bin) {
    int cur = 0;
    for (int i = 0; i < nd - 1; ++i) {
      cur += counts[i];
      if (cur >= min_data_in_bin) {
        double val = next_up((dv[i] + dv[i + 1]) / 2.0);
        if (ub.empty() || !check_double_equal_ordered(ub.back(), val)) {
          ub.push_back(val);
          cur = 0;
        }
      }
    }
    ub.push_back(INFINITY);
  } else {
    if (min_data_in_bin > 0) {
      max_bin = std::min(max_bin, (int)(total_cnt / (size_t)min_data_in_bin));
      max_bin = std::max(max_bin, 1);
    }
    double mean_bin_size = (double)total_cnt / (double)max_bin;
    int rest_bin_cnt = max_bin;
    int rest_sample_cnt = (int)total_cnt;
    std::vector<char> big(nd, 0);
    for (int i = 0; i < nd; ++i) {
      if ((double)counts[i] >= mean_bin_size) {
        big[i] = 1;
        rest_bin_cnt -= 1;
        rest_sample_cnt -= counts[i];
      }
    }
    mean_bin_size = (double)rest_sample_cnt / (double)rest_bin_cnt;
    std::vector<double> upper(max_bin, INFINITY), lower(max_bin, INFINITY);
    int bin_cnt = 0;
    lower[0] = dv[0];
    int cur = 0;
    for (int i = 0; i < nd - 1; ++i) {
      if (!big[i]) rest_sample_cnt -= counts[i];
      cur += counts[i];
      double half = std::max(mean_bin_size * (double)0.5f, 1.0);
      if (big[i] || (double)cur >= mean_bin_size || (big[i + 1] && (double)cur >= half)) {
        upper[bin_cnt] = dv[i];
        bin_cnt += 1;
        lower[bin_cnt] = dv[i + 1];
        if (bin_cnt >= max_bin - 1) break;
        cur = 0;
        if (!big[i]) {
          rest_bin_cnt -= 1;
          mean_bin_size = (double)rest_sample_cnt / (double)rest_bin_cnt;
        }
      }
    }
    bin_cnt += 1;
    for (int i = 0; i < bin_cnt - 1; ++i) {
      double val = next_up((upper[i] + lower[i + 1]) / 2.0);
      if (ub.empty() || !check_double_equal_ordered(ub.back(), val)) ub.push_back(val);
    }
    ub.push_back(INFINITY);
  }
  return ub;
}

// binning.rs:215-276 (no forced bounds)
static std::vector<double> find_bin_with_zero_as_one_bin(const std::vector<double>& dv, const std::vector<int>& counts,
                                                         int max_bin, size_t total_sample_cnt, int min_data_in_bin) {
  size_t n = dv.size();
  std::vector<double> ub;
  int left_cnt_data = 0, cnt_zero = 0, right_cnt_data = 0;
  for (size_t i = 0; i < n; ++i) {
    if (dv[i] <= -K_ZERO_THRESHOLD) left_cnt_data += counts[i];
    else if (dv[i] > K_ZERO_THRESHOLD) right_cnt_data += counts[i];
    else cnt_zero += counts[i];
  }
  size_t left_cnt = n;
  for (size_t i = 0; i < n; ++i) if (dv[i] > -K_ZERO_THRESHOLD) { left_cnt = i; break; }
  if (left_cnt > 0 && max_bin > 1) {
    double denom = (double)(total_sample_cnt - (size_t)cnt_zero);
    int left_max_bin = (int)((double)left_cnt_data / denom * (double)(max_bin - 1));
    left_max_bin = std::max(left_max_bin, 1);
    ub = greedy_find_bin(dv.data(), counts.data(), (int)left_cnt, left_max_bin, (size_t)left_cnt_data, min_data_in_bin);
    if (!ub.empty()) ub.back() = -K_ZERO_THRESHOLD;
  }
  long right_start = -1;
  for (size_t i = left_cnt; i < n; ++i) if (dv[i] > K_ZERO_THRESHOLD) { right_start = (long)i; break; }
  int right_max_bin = max_bin - 1 - (int)ub.size();
  if (right_start >= 0 && right_max_bin > 0) {
    auto rb = greedy_find_bin(dv.data() + right_start, counts.data() + right_start, (int)(n - right_start),
                              right_max_bin, (size_t)right_cnt_data, min_data_in_bin);
    ub.push_back(K_ZERO_THRESHOLD);
    ub.insert(ub.end(), rb.begin(), rb.end());
  } else {
    ub.push_back(INFINITY);
  }
  return ub;
}

// binning.rs:480-517 (finish) and binning.rs:574-602
static void finish_mapper(BinMapper& m, const std::vector<int>& cnt_in_bin, size_t total_sample_cnt, const BinParams& p) {
  m.is_trivial = m.num_bin <= 1;
  if (!m.is_trivial && p.pre_filter && need_filter(cnt_in_bin, (int)total_sample_cnt, p.min_split_data, m.bin_type))
    m.is_trivial = true;
  if (!m.is_trivial) {
    m.default_bin = m.value_to_bin(0.0);
    size_t arg_max = 0;
    for (size_t i = 1; i < cnt_in_bin.size(); ++i) if (cnt_in_bin[i] > cnt_in_bin[arg_max]) arg_max = i;
    m.most_freq_bin = (uint32_t)arg_max;
    double max_sparse_rate = (double)cnt_in_bin[arg_max] / (double)total_sample_cnt;
    if (m.most_freq_bin != m.default_bin && max_sparse_rate < K_SPARSE_THRESHOLD) m.most_freq_bin = m.default_bin;
    m.sparse_rate = (double)cnt_in_bin[m.most_freq_bin] / (double)total_sample_cnt;
  } else {
    m.sparse_rate = 1.0;
  }
}

// binning.rs:395-472 (numerical) and 476-573 (categorical)
static BinMapper find_bin(const std::vector<double>& values, size_t total_sample_cnt, int bin_type, const BinParams& p) {
  int num_in = (int)values.size();
  std::vector<double> vals;
  for (double v : values) if (!std::isnan(v)) vals.push_back(v);
  int non_na = (int)vals.size();
  int na_cnt = 0;
  int missing_type;
  if (!p.use_missing) missing_type = MT_NONE;
  else if (p.zero_as_missing) missing_type = MT_ZERO;
  else if (non_na == num_in) missing_type = MT_NONE;
  else { na_cnt = num_in - non_na; missing_type = MT_NAN; }
  int nsv = non_na;
  int zero_cnt = (int)((int64_t)total_sample_cnt - nsv - na_cnt);
  std::stable_sort(vals.begin(), vals.end());
  std::vector<double> dv;
  std::vector<int> counts;
  if (nsv == 0 || (vals[0] > 0.0 && zero_cnt > 0)) { dv.push_back(0.0); counts.push_back(zero_cnt); }
  if (nsv > 0) { dv.push_back(vals[0]); counts.push_back(1); }
  for (int i = 1; i < nsv; ++i) {
    if (!check_double_equal_ordered(vals[i - 1], vals[i])) {
      if (vals[i - 1] < 0.0 && vals[i] > 0.0) { dv.push_back(0.0); counts.push_back(zero_cnt); }
      dv.push_back(vals[i]);
      counts.push_back(1);
    } else {
      dv.back() = vals[i];
      counts.back() += 1;
    }
  }
  if (nsv > 0 && vals[nsv - 1] < 0.0 && zero_cnt > 0) { dv.push_back(0.0); counts.push_back(zero_cnt); }
  BinMapper m;
  m.min_val = dv[0];
  m.max_val = dv.back();
  m.missing_type = missing_type;
  m.bin_type = bin_type;
  if (bin_type == BT_CAT) {
    m.bin_upper_bound.clear();
    std::vector<int> dvi, ci;
    for (size_t i = 0; i < dv.size(); ++i) {
      int val = (int)dv[i];
      if (val < 0) na_cnt += counts[i];
      else if (dvi.empty() || dvi.back() != val) { dvi.push_back(val); ci.push_back(counts[i]); }
      else ci.back() += counts[i];
    }
    std::vector<int> cnt_in_bin;
    int rest_cnt = (int)((int64_t)total_sample_cnt - na_cnt);
    if (rest_cnt > 0) {
      std::vector<std::pair<int, int>> pairs;
      for (size_t i = 0; i < dvi.size(); ++i) pairs.push_back({ci[i], dvi[i]});
      std::stable_sort(pairs.begin(), pairs.end(), [](const std::pair<int, int>& a, const std::pair<int, int>& b) {
        return a.first > b.first;
      });
      int cut_cnt = (int)((double)((float)(total_sample_cnt - (size_t)na_cnt) * 0.99f) + (double)0.5f);
      int distinct_cnt = (int)pairs.size();
      if (na_cnt > 0) distinct_cnt += 1;
      int max_bin = std::min(distinct_cnt, p.max_bin);
      m.bin_2_categorical.push_back(-1);
      m.categorical_2_bin[-1] = 0;
      cnt_in_bin.push_back(0);
      m.num_bin = 1;
      int used_cnt = 0;
      size_t cur = 0;
      while (cur < pairs.size() && (used_cnt < cut_cnt || m.num_bin < max_bin)) {
        if (pairs[cur].first < p.min_data_in_bin && cur > 1) break;
        m.bin_2_categorical.push_back(pairs[cur].second);
        m.categorical_2_bin[pairs[cur].second] = (uint32_t)m.num_bin;
        used_cnt += pairs[cur].first;
        cnt_in_bin.push_back(pairs[cur].first);
        m.num_bin += 1;
        cur += 1;
      }
      m.missing_type = (cur == pairs.size() && na_cnt == 0) ? MT_NONE : MT_NAN;
      cnt_in_bin[0] = (int)((int64_t)total_sample_cnt - used_cnt);
    }
    finish_mapper(m, cnt_in_bin, total_sample_cnt, p);
    return m;
  }
  std::vector<double> ub;
  if (missing_type == MT_ZERO) {
    ub = find_bin_with_zero_as_one_bin(dv, counts, p.max_bin, total_sample_cnt, p.min_data_in_bin);
    if (ub.size() == 2) missing_type = MT_NONE;
  } else if (missing_type == MT_NONE) {
    ub = find_bin_with_zero_as_one_bin(dv, counts, p.max_bin, total_sample_cnt, p.min_data_in_bin);
  } else {
    ub = find_bin_with_zero_as_one_bin(dv, counts, p.max_bin - 1, total_sample_cnt - (size_t)na_cnt, p.min_data_in_bin);
    ub.push_back(NAN);
  }
  m.missing_type = missing_type;
  int num_bin = (int)ub.size();
  std::vector<int> cnt_in_bin(num_bin, 0);
  size_t ib = 0;
  for (size_t i = 0; i < dv.size(); ++i) {
    while (dv[i] > ub[ib] && (int)ib < num_bin - 1) ib++;
    cnt_in_bin[ib] += counts[i];
  }
  if (missing_type == MT_NAN) cnt_in_bin[num_bin - 1] = na_cnt;
  m.num_bin = num_bin;
  m.bin_upper_bound = ub;
  finish_mapper(m, cnt_in_bin, total_sample_cnt, p);
  return m;
}

// ---------------------------------------------------------------- dataset
struct Dataset {
  int64_t n = 0;
  int ncol = 0;
  std::vector<BinMapper> mappers;   // per column
  std::vector<int> used;            // inner -> real column
  std::vector<int> real_to_inner;   // real -> inner or -1
  std::vector<uint8_t> bins;        // n x F (inner), raw bin per feature (u8)
  std::vector<uint16_t> bins16;     // used instead when some num_bin > 256
  bool wide = false;
  std::vector<float> label;
  std::vector<int> query_boundaries;
  int num_groups = 0;               // upstream's feature groups (bundling), for the bagging subset rule
  uint32_t bin(int64_t row, int f) const {
    size_t F = used.size();
    return wide ? bins16[(size_t)row * F + f] : bins[(size_t)row * F + f];
  }
};

struct Params {
  bool enable_bundle = true;   // dataset.rs:835-853 (explicit enable_bundle, default true)
  int max_bin = 255, min_data_in_bin = 3, bin_construct_sample_cnt = 200000, data_random_seed = 1;
  int min_data_in_leaf = 20;
  double min_sum_hessian_in_leaf = 1e-3;
  bool feature_pre_filter = true, use_missing = true, zero_as_missing = false;
  double lambda_l1 = 0, lambda_l2 = 0, max_delta_step = 0, path_smooth = 0, min_gain_to_split = 0;
  int max_cat_to_onehot = 4, max_cat_threshold = 32;
  double cat_l2 = 10.0, cat_smooth = 10.0;
  int min_data_per_group = 100;
  int num_leaves = 31, max_depth = -1;
  double learning_rate = 0.1;
  int objective = 0;  // 0 binary, 1 multiclass, 2 lambdarank
  int num_class = 1;
  double sigmoid = 1.0;
  bool boost_from_average = true;
  bool use_quantized_grad = false;
  int num_grad_quant_bins = 4;
  bool stochastic_rounding = true, quant_train_renew_leaf = false;
  int seed = 0;
  int lambdarank_truncation_level = 30;
  bool lambdarank_norm = true;
  int num_threads = 1;
  double bagging_fraction = 1.0;   // sample_strategy.rs (row bagging)
  int bagging_freq = 0, bagging_seed = 3;
  double pos_bagging_fraction = 1.0, neg_bagging_fraction = 1.0;
  int goss = 0;                    // data_sample_strategy=goss
  double top_rate = 0.2, other_rate = 0.1;
  bool bagging_by_query = false;
  std::vector<int> categorical;
};

static bool parse_params(const char* s, Params& p) {
  std::string str(s ? s : "");
  size_t i = 0;
  bool seed_given = false, drs_given = false, bs_given = false;
  while (i < str.size()) {
    while (i < str.size() && (str[i] == ' ' || str[i] == ',')) ++i;
    size_t j = i;
    while (j < str.size() && str[j] != ' ' && str[j] != ',') ++j;
    if (j > i) {
      std::string kv = str.substr(i, j - i);
      size_t eq = kv.find('=');
      if (eq == std::string::npos) return false;
      std::string k = kv.substr(0, eq), v = kv.substr(eq + 1);
      if (k == "max_bin") p.max_bin = atoi(v.c_str());
      else if (k == "min_data_in_bin") p.min_data_in_bin = atoi(v.c_str());
      else if (k == "bin_construct_sample_cnt") p.bin_construct_sample_cnt = atoi(v.c_str());
      else if (k == "data_random_seed") p.data_random_seed = atoi(v.c_str());
      else if (k == "min_data_in_leaf") p.min_data_in_leaf = atoi(v.c_str());
      else if (k == "min_sum_hessian_in_leaf") p.min_sum_hessian_in_leaf = atof(v.c_str());
      else if (k == "feature_pre_filter") p.feature_pre_filter = v == "true" || v == "1";
      else if (k == "enable_bundle") p.enable_bundle = v == "true" || v == "1";
      else if (k == "use_missing") p.use_missing = v == "true" || v == "1";
      else if (k == "zero_as_missing") p.zero_as_missing = v == "true" || v == "1";
      else if (k == "lambda_l1") p.lambda_l1 = atof(v.c_str());
      else if (k == "lambda_l2") p.lambda_l2 = atof(v.c_str());
      else if (k == "max_delta_step") p.max_delta_step = atof(v.c_str());
      else if (k == "path_smooth") p.path_smooth = atof(v.c_str());
      else if (k == "min_gain_to_split") p.min_gain_to_split = atof(v.c_str());
      else if (k == "max_cat_to_onehot") p.max_cat_to_onehot = atoi(v.c_str());
      else if (k == "max_cat_threshold") p.max_cat_threshold = atoi(v.c_str());
      else if (k == "cat_l2") p.cat_l2 = atof(v.c_str());
      else if (k == "cat_smooth") p.cat_smooth = atof(v.c_str());
      else if (k == "min_data_per_group") p.min_data_per_group = atoi(v.c_str());
      else if (k == "num_leaves") p.num_leaves = atoi(v.c_str());
      else if (k == "max_depth") p.max_depth = atoi(v.c_str());
      else if (k == "learning_rate") p.learning_rate = atof(v.c_str());
      else if (k == "objective") p.objective = v == "binary" ? 0 : v == "multiclass" ? 1 : v == "lambdarank" ? 2 : -1;
      else if (k == "num_class") p.num_class = atoi(v.c_str());
      else if (k == "sigmoid") p.sigmoid = atof(v.c_str());
      else if (k == "boost_from_average") p.boost_from_average = v == "true" || v == "1";
      else if (k == "use_quantized_grad") p.use_quantized_grad = v == "true" || v == "1";
      else if (k == "num_grad_quant_bins") p.num_grad_quant_bins = atoi(v.c_str());
      else if (k == "stochastic_rounding") p.stochastic_rounding = v == "true" || v == "1";
      else if (k == "quant_train_renew_leaf") p.quant_train_renew_leaf = v == "true" || v == "1";
      else if (k == "seed") p.seed = atoi(v.c_str());
      else if (k == "lambdarank_truncation_level") p.lambdarank_truncation_level = atoi(v.c_str());
      else if (k == "lambdarank_norm") p.lambdarank_norm = v == "true" || v == "1";
      else if (k == "num_threads") p.num_threads = atoi(v.c_str());
      else if (k == "bagging_fraction") p.bagging_fraction = atof(v.c_str());
      else if (k == "bagging_freq") p.bagging_freq = atoi(v.c_str());
      else if (k == "bagging_seed") p.bagging_seed = atoi(v.c_str());
      else if (k == "pos_bagging_fraction") p.pos_bagging_fraction = atof(v.c_str());
      else if (k == "data_sample_strategy") p.goss = v == "goss";
      else if (k == "top_rate") p.top_rate = atof(v.c_str());
      else if (k == "other_rate") p.other_rate = atof(v.c_str());
      else if (k == "neg_bagging_fraction") p.neg_bagging_fraction = atof(v.c_str());
      else if (k == "bagging_by_query") p.bagging_by_query = v == "true" || v == "1";
      else if (k == "categorical_feature") {
        size_t a = 0;
        while (a < v.size()) {
          size_t b = v.find(':', a);
          if (b == std::string::npos) b = v.size();
          p.categorical.push_back(atoi(v.substr(a, b - a).c_str()));
          a = b + 1;
        }
      } else if (k == "num_iterations" || k == "force_row_wise" || k == "deterministic" || k == "verbose") {
      } else return false;
      if (p.objective < 0) return false;
      seed_given |= k == "seed";
      drs_given |= k == "data_random_seed";
      bs_given |= k == "bagging_seed";
    }
    i = j;
  }
  if (seed_given) {
    // config.rs:795-806: an explicit seed derives the other seeds; explicit
    // values of those still win
    Random r(p.seed);
    const int drs = r.next_short(0, 32767), bs = r.next_short(0, 32767);
    if (!drs_given) p.data_random_seed = drs;
    if (!bs_given) p.bagging_seed = bs;
  }
  return true;
}

// Upstream's feature-group count (feature_groups.rs:169-211 upstream_groups,
// is_enable_sparse on): FindGroups in two feature orders, the fewer groups
// kept. Only the count is used (the bagging subset rule); the bins stay per
// feature. cols[f]: the sample positions of used feature f's kept values
// (matrix.rs:198-214), ascending; vals[f] those values.
static int bundle_group_count(const Dataset* ds, const std::vector<std::vector<int>>& cols,
                              const std::vector<std::vector<double>>& vals, int64_t total) {
  const int F = (int)ds->used.size();
  if (F == 0 || total == 0) return F;
  // FixSampleIndices (feature_groups.rs:21-42)
  std::vector<std::vector<int>> si(F);
  for (int f = 0; f < F; ++f) {
    const BinMapper& m = ds->mappers[ds->used[f]];
    std::vector<int> fixed;
    if (m.default_bin != m.most_freq_bin) {
      int j = 0;
      const int nv = (int)cols[f].size();
      for (int i = 0; i < (int)total;) {
        if (j < nv && cols[f][j] < i) {
          ++j;
        } else if (j < nv && cols[f][j] == i) {
          if (m.value_to_bin(vals[f][j]) != m.most_freq_bin) fixed.push_back(i);
          ++i;
        } else {
          fixed.push_back(i);
          ++i;
        }
      }
    }
    si[f] = fixed.empty() ? cols[f] : fixed;
  }
  // FindGroups (feature_groups.rs:62-158), counting the groups
  auto find_groups = [&](const std::vector<int>& order) {
    const int64_t max_cnt = total / 10000;
    Random rand((int)ds->n);
    std::vector<std::vector<bool>> marks;
    std::vector<int64_t> total_cnt, used_cnt;
    for (int fidx : order) {
      const std::vector<int>& idx = si[fidx];
      const int64_t cur = (int64_t)idx.size();
      std::vector<int> available;
      for (int g = 0; g < (int)marks.size(); ++g)
        if (total_cnt[g] + cur <= total + max_cnt) available.push_back(g);
      std::vector<int> search;
      if (!available.empty()) {
        const int last = (int)available.size() - 1;
        std::vector<int> picked = rand.sample(last, std::min(last, 100 - 1));
        search.push_back(available.back());
        for (int k : picked) search.push_back(available[k]);
      }
      int best = -1;
      int64_t best_cnt = 0;
      for (int g : search) {
        const int64_t rest = max_cnt - total_cnt[g] + used_cnt[g];
        int64_t cnt = 0;   // GetConflictCount: -1 once past rest
        for (int i : idx) {
          if (marks[g][i]) ++cnt;
          if (cnt > rest) {
            cnt = -1;
            break;
          }
        }
        if (cnt >= 0 && cnt <= rest && cnt <= cur / 2) {
          best = g;
          best_cnt = cnt;
          break;
        }
      }
      if (best >= 0) {
        total_cnt[best] += cur;
        used_cnt[best] += cur - best_cnt;
        for (int i : idx) marks[best][i] = true;
      } else {
        marks.emplace_back((size_t)total, false);
        for (int i : idx) marks.back()[i] = true;
        total_cnt.push_back(cur);
        used_cnt.push_back(cur);
      }
    }
    // the second round: groups under the dense threshold merge into one
    int kept = 0, sparse = 0;
    for (size_t g = 0; g < marks.size(); ++g) {
      if ((double)used_cnt[g] / (double)total >= 0.4) ++kept;
      else ++sparse;
    }
    return kept + (sparse > 0 ? 1 : 0);
  };
  std::vector<int> in_order(F), by_cnt(F);
  for (int f = 0; f < F; ++f) in_order[f] = by_cnt[f] = f;
  std::stable_sort(by_cnt.begin(), by_cnt.end(),
                   [&](int a, int b) { return cols[b].size() < cols[a].size(); });
  const int g1 = find_groups(in_order), g2 = find_groups(by_cnt);
  return g1 > g2 ? g2 : g1;
}

// dataset.rs:364-422 with the dense-matrix sampling of matrix.rs:198-214.
// Exclusive feature bundling changes only the storage of the bins: every
// used feature keeps its own bins (value_to_bin of its value), and the
// number of groups it would form is restated above for the bagging rule.
static Dataset* build_dataset(const float* X, int64_t n, int ncol, const Params& p) {
  Dataset* ds = new Dataset();
  ds->n = n;
  ds->ncol = ncol;
  int sample_cnt = (int)std::min<int64_t>(n, std::max(p.bin_construct_sample_cnt, 0));
  Random rand(p.data_random_seed);
  std::vector<int> si = rand.sample((int)n, sample_cnt);
  size_t total = si.size();
  int filter_cnt = (int)((double)p.min_data_in_leaf * (double)total / (double)n);
  std::vector<char> is_cat(ncol, 0);
  for (int c : p.categorical) if (c >= 0 && c < ncol) is_cat[c] = 1;
  for (int c = 0; c < ncol; ++c) {
    std::vector<double> vals;
    for (size_t i = 0; i < total; ++i) {
      double v = (double)X[(size_t)si[i] * ncol + c];
      if (std::fabs(v) > K_ZERO_THRESHOLD || std::isnan(v)) vals.push_back(v);
    }
    BinParams bp{p.max_bin, p.min_data_in_bin, filter_cnt, p.feature_pre_filter, p.use_missing, p.zero_as_missing};
    ds->mappers.push_back(find_bin(vals, total, is_cat[c] ? BT_CAT : BT_NUM, bp));
  }
  ds->real_to_inner.assign(ncol, -1);
  for (int c = 0; c < ncol; ++c) {
    if (!ds->mappers[c].is_trivial) {
      ds->real_to_inner[c] = (int)ds->used.size();
      ds->used.push_back(c);
      if (ds->mappers[c].num_bin > 256) ds->wide = true;
    }
  }
  size_t F = ds->used.size();
  if (p.enable_bundle) {
    std::vector<std::vector<int>> cols(F);
    std::vector<std::vector<double>> cvals(F);
    for (size_t f = 0; f < F; ++f) {
      for (size_t i = 0; i < total; ++i) {
        double v = (double)X[(size_t)si[i] * ncol + ds->used[f]];
        if (std::fabs(v) > K_ZERO_THRESHOLD || std::isnan(v)) {
          cols[f].push_back((int)i);
          cvals[f].push_back(v);
        }
      }
    }
    ds->num_groups = bundle_group_count(ds, cols, cvals, (int64_t)total);
  } else {
    ds->num_groups = (int)F;
  }
  if (ds->wide) ds->bins16.resize((size_t)n * F); else ds->bins.resize((size_t)n * F);
  for (int64_t r = 0; r < n; ++r) {
    for (size_t f = 0; f < F; ++f) {
      uint32_t b = ds->mappers[ds->used[f]].value_to_bin((double)X[(size_t)r * ncol + ds->used[f]]);
      if (ds->wide) ds->bins16[(size_t)r * F + f] = (uint16_t)b; else ds->bins[(size_t)r * F + f] = (uint8_t)b;
    }
  }
  return ds;
}

// ---------------------------------------------------------------- split search
// split.rs:18-29
struct FeatureMeta {
  int num_bin, missing_type, offset;
  uint32_t default_bin, most_freq_bin;
  int bin_type;
  double penalty = 1.0;
};

// split.rs:33-46
struct SplitParams {
  double lambda_l1, lambda_l2, max_delta_step, path_smooth;
  int min_data_in_leaf;
  double min_sum_hessian_in_leaf, min_gain_to_split;
  int max_cat_to_onehot, max_cat_threshold;
  double cat_l2, cat_smooth;
  int min_data_per_group;
  bool use_l1() const { return lambda_l1 > 0.0; }
  bool use_max_output() const { return max_delta_step > 0.0; }
  bool use_smoothing() const { return path_smooth > K_EPSILON; }
};

// split.rs:60-122
struct SplitInfo {
  int feature = -1;
  uint32_t threshold = 0;
  std::vector<uint32_t> cat_threshold;
  double left_output = 0, right_output = 0, gain = K_MIN_SCORE;
  int left_count = 0, right_count = 0;
  double left_sum_gradient = 0, left_sum_hessian = 0, right_sum_gradient = 0, right_sum_hessian = 0;
  bool default_left = true;
  int64_t left_sum_gh = 0, right_sum_gh = 0;
  void reset() { feature = -1; gain = K_MIN_SCORE; }
  bool better_than(const SplitInfo& o) const {
    if (gain != o.gain) return gain > o.gain;
    int a = feature == -1 ? INT32_MAX : feature;
    int b = o.feature == -1 ? INT32_MAX : o.feature;
    return a < b;
  }
};

static inline double sgn(double x) { return (double)((x > 0.0) - (x < 0.0)); }
static inline double threshold_l1(double s, double l1) { return sgn(s) * std::max(std::fabs(s) - l1, 0.0); }

// split.rs:137-161
static double leaf_output_t(double g, double h, const SplitParams& p, bool use_l1, bool use_max, bool use_smooth,
                            int num_data, double parent_output) {
  double ret = use_l1 ? -threshold_l1(g, p.lambda_l1) / (h + p.lambda_l2) : -g / (h + p.lambda_l2);
  if (use_max && p.max_delta_step > 0.0 && std::fabs(ret) > p.max_delta_step) ret = sgn(ret) * p.max_delta_step;
  if (use_smooth) {
    double nn = (double)num_data;
    ret = ret * (nn / p.path_smooth) / (nn / p.path_smooth + 1.0) + parent_output / (nn / p.path_smooth + 1.0);
  }
  return ret;
}
static double leaf_output(double g, double h, const SplitParams& p, int nd, double po) {
  return leaf_output_t(g, h, p, p.use_l1(), p.use_max_output(), p.use_smoothing(), nd, po);
}
static double refit_leaf_output(double g, double h, const SplitParams& p, bool smooth, int nd, double po) {
  return leaf_output_t(g, h, p, true, true, smooth, nd, po);
}
static double root_output(double g, double h, const SplitParams& p, int nd) {
  return leaf_output_t(g, h, p, true, true, false, nd, 0.0);
}
// split.rs:183-201
static double leaf_gain_given_output(double g, double h, const SplitParams& p, double output) {
  double sg = p.use_l1() ? threshold_l1(g, p.lambda_l1) : g;
  return -(2.0 * sg * output + (h + p.lambda_l2) * output * output);
}
static double leaf_gain(double g, double h, const SplitParams& p, int nd, double po) {
  if (!p.use_max_output() && !p.use_smoothing()) {
    if (p.use_l1()) { double sg = threshold_l1(g, p.lambda_l1); return (sg * sg) / (h + p.lambda_l2); }
    return (g * g) / (h + p.lambda_l2);
  }
  double out = leaf_output(g, h, p, nd, po);
  return leaf_gain_given_output(g, h, p, out);
}
// split.rs:203-226 (no monotone constraints)
static double split_gain(double lg, double lh, double rg, double rh, const SplitParams& p, int lc, int rc, double po) {
  return leaf_gain(lg, lh, p, lc, po) + leaf_gain(rg, rh, p, rc, po);
}
// split.rs:228-231
static inline int round_int(double x) { return (int)(x + (double)0.5f); }

// split.rs:676-891 (FindBestThresholdSequentially; no monotone constraints, no extra trees).
// `hist` is the per-feature view: [g, h] pairs for stored bins t = bin - offset.
static bool scan(const double* hist, const FeatureMeta& meta, const SplitParams& p, double sum_gradient,
                 double sum_hessian, int num_data, double min_gain_shift, double parent_output, bool reverse,
                 bool skip_default_bin, bool na_as_missing, SplitInfo& out) {
  int offset = meta.offset;
  auto g = [&](int t) { return hist[2 * t]; };
  auto h = [&](int t) { return hist[2 * t + 1]; };
  bool is_splittable = false;
  double best_slg = NAN, best_slh = NAN, best_gain = K_MIN_SCORE;
  int best_lc = 0;
  uint32_t best_threshold = (uint32_t)meta.num_bin;
  double cnt_factor = (double)num_data / sum_hessian;
  int min_data = p.min_data_in_leaf;
  double min_hess = p.min_sum_hessian_in_leaf;
  if (reverse) {
    double srg = 0.0, srh = K_EPSILON;
    int rc = 0;
    int t = meta.num_bin - 1 - offset - (int)na_as_missing;
    int t_end = 1 - offset;
    while (t >= t_end) {
      if (skip_default_bin && (t + offset) == (int)meta.default_bin) { t--; continue; }
      double gr = g(t), hs = h(t);
      int cnt = round_int(hs * cnt_factor);
      srg += gr; srh += hs; rc += cnt;
      if (rc < min_data || srh < min_hess) { t--; continue; }
      int lc = num_data - rc;
      if (lc < min_data) break;
      double slh = sum_hessian - srh;
      if (slh < min_hess) break;
      double slg = sum_gradient - srg;
      double cg = split_gain(slg, slh, srg, srh, p, lc, rc, parent_output);
      if (cg <= min_gain_shift) { t--; continue; }
      is_splittable = true;
      if (cg > best_gain) {
        best_lc = lc; best_slg = slg; best_slh = slh;
        best_threshold = (uint32_t)(t - 1 + offset);
        best_gain = cg;
      }
      t--;
    }
  } else {
    double slg = 0.0, slh = K_EPSILON;
    int lc = 0;
    int t = 0;
    int t_end = meta.num_bin - 2 - offset;
    if (na_as_missing && offset == 1) {
      slg = sum_gradient;
      slh = sum_hessian - K_EPSILON;
      lc = num_data;
      for (int i = 0; i < meta.num_bin - offset; ++i) {
        slg -= g(i); slh -= h(i); lc -= round_int(h(i) * cnt_factor);
      }
      t = -1;
    }
    while (t <= t_end) {
      if (skip_default_bin && (t + offset) == (int)meta.default_bin) { t++; continue; }
      if (t >= 0) { slg += g(t); slh += h(t); lc += round_int(h(t) * cnt_factor); }
      if (lc < min_data || slh < min_hess) { t++; continue; }
      int rc = num_data - lc;
      if (rc < min_data) break;
      double srh = sum_hessian - slh;
      if (srh < min_hess) break;
      double srg = sum_gradient - slg;
      double cg = split_gain(slg, slh, srg, srh, p, lc, rc, parent_output);
      if (cg <= min_gain_shift) { t++; continue; }
      is_splittable = true;
      if (cg > best_gain) {
        best_lc = lc; best_slg = slg; best_slh = slh;
        best_threshold = (uint32_t)(t + offset);
        best_gain = cg;
      }
      t++;
    }
  }
  if (is_splittable && best_gain > out.gain + min_gain_shift) {
    out.threshold = best_threshold;
    out.left_output = leaf_output(best_slg, best_slh, p, best_lc, parent_output);
    out.left_count = best_lc;
    out.left_sum_gradient = best_slg;
    out.left_sum_hessian = best_slh - K_EPSILON;
    out.right_output = leaf_output(sum_gradient - best_slg, sum_hessian - best_slh, p, num_data - best_lc, parent_output);
    out.right_count = num_data - best_lc;
    out.right_sum_gradient = sum_gradient - best_slg;
    out.right_sum_hessian = sum_hessian - best_slh - K_EPSILON;
    out.gain = best_gain - min_gain_shift;
    out.default_left = reverse;
  }
  return is_splittable;
}

// split.rs:402-441
static bool find_best_threshold_numerical(const double* hist, const FeatureMeta& meta, const SplitParams& p,
                                          double sg, double sh, int nd, double po, SplitInfo& out) {
  double min_gain_shift = leaf_gain(sg, sh, p, nd, po) + p.min_gain_to_split;
  bool splittable = false;
  if (meta.num_bin > 2 && meta.missing_type != MT_NONE) {
    if (meta.missing_type == MT_ZERO) {
      splittable |= scan(hist, meta, p, sg, sh, nd, min_gain_shift, po, true, true, false, out);
      splittable |= scan(hist, meta, p, sg, sh, nd, min_gain_shift, po, false, true, false, out);
    } else {
      splittable |= scan(hist, meta, p, sg, sh, nd, min_gain_shift, po, true, false, true, out);
      splittable |= scan(hist, meta, p, sg, sh, nd, min_gain_shift, po, false, false, true, out);
    }
  } else {
    splittable |= scan(hist, meta, p, sg, sh, nd, min_gain_shift, po, true, false, false, out);
    if (meta.missing_type == MT_NAN) out.default_left = false;
  }
  return splittable;
}

// split.rs:446-670 (FindBestThresholdCategoricalInner, no monotone constraints)
static bool find_best_threshold_categorical(const double* hist, const FeatureMeta& meta, const SplitParams& p,
                                            double sum_gradient, double sum_hessian, int num_data,
                                            double parent_output, SplitInfo& out) {
  auto g = [&](int t) { return hist[2 * t]; };
  auto h = [&](int t) { return hist[2 * t + 1]; };
  bool is_splittable = false;
  out.default_left = false;
  double best_gain = K_MIN_SCORE, best_slg = 0, best_slh = 0;
  int best_lc = 0;
  double gain_shift = p.use_smoothing() ? leaf_gain_given_output(sum_gradient, sum_hessian, p, parent_output)
                                        : leaf_gain(sum_gradient, sum_hessian, p, num_data, 0.0);
  double min_gain_shift = gain_shift + p.min_gain_to_split;
  int offset = meta.offset;
  int bin_start = 1 - offset, bin_end = meta.num_bin - offset;
  int used_bin = -1;
  std::vector<int> sorted_idx;
  SplitParams pc = p;
  bool use_onehot = meta.num_bin <= p.max_cat_to_onehot;
  int best_threshold = -1, best_dir = 1;
  double cnt_factor = (double)num_data / sum_hessian;
  if (use_onehot) {
    for (int t = bin_start; t < bin_end; ++t) {
      double grad = g(t), hess = h(t);
      int cnt = round_int(hess * cnt_factor);
      if (cnt < p.min_data_in_leaf || hess < p.min_sum_hessian_in_leaf) continue;
      int other_count = num_data - cnt;
      if (other_count < p.min_data_in_leaf) continue;
      double soh = sum_hessian - hess - K_EPSILON;
      if (soh < p.min_sum_hessian_in_leaf) continue;
      double sog = sum_gradient - grad;
      double cg = split_gain(sog, soh, grad, hess + K_EPSILON, pc, other_count, cnt, parent_output);
      if (cg <= min_gain_shift) continue;
      is_splittable = true;
      if (cg > best_gain) {
        best_threshold = t; best_slg = grad; best_slh = hess + K_EPSILON; best_lc = cnt; best_gain = cg;
      }
    }
  } else {
    for (int i = bin_start; i < bin_end; ++i)
      if ((double)round_int(h(i) * cnt_factor) >= p.cat_smooth) sorted_idx.push_back(i);
    used_bin = (int)sorted_idx.size();
    pc.lambda_l2 += p.cat_l2;
    auto ctr = [&](int i) { return g(i) / (h(i) + p.cat_smooth); };
    std::stable_sort(sorted_idx.begin(), sorted_idx.end(), [&](int i, int j) { return ctr(i) < ctr(j); });
    int max_num_cat = std::min(p.max_cat_threshold, (used_bin + 1) / 2);
    is_splittable = false;
    const int dirs[2] = {1, -1};
    for (int d = 0; d < 2; ++d) {
      int dir = dirs[d];
      int start_pos = dir == 1 ? 0 : used_bin - 1;
      int cnt_cur_group = 0;
      double slg = 0.0, slh = K_EPSILON;
      int lc = 0;
      int i = 0;
      while (i < used_bin && i < max_num_cat) {
        int t = sorted_idx[start_pos];
        start_pos += dir;
        double grad = g(t), hess = h(t);
        int cnt = round_int(hess * cnt_factor);
        slg += grad; slh += hess; lc += cnt; cnt_cur_group += cnt;
        if (lc < p.min_data_in_leaf || slh < p.min_sum_hessian_in_leaf) { i++; continue; }
        int rc = num_data - lc;
        if (rc < p.min_data_in_leaf || rc < p.min_data_per_group) break;
        double srh = sum_hessian - slh;
        if (srh < p.min_sum_hessian_in_leaf) break;
        if (cnt_cur_group < p.min_data_per_group) { i++; continue; }
        cnt_cur_group = 0;
        double srg = sum_gradient - slg;
        double cg = split_gain(slg, slh, srg, srh, pc, lc, rc, parent_output);
        if (cg <= min_gain_shift) { i++; continue; }
        is_splittable = true;
        if (cg > best_gain) {
          best_lc = lc; best_slg = slg; best_slh = slh; best_threshold = i; best_gain = cg; best_dir = dir;
        }
        i++;
      }
    }
  }
  if (is_splittable) {
    out.left_output = leaf_output(best_slg, best_slh, pc, best_lc, parent_output);
    out.left_count = best_lc;
    out.left_sum_gradient = best_slg;
    out.left_sum_hessian = best_slh - K_EPSILON;
    out.right_output = leaf_output(sum_gradient - best_slg, sum_hessian - best_slh, pc, num_data - best_lc, parent_output);
    out.right_count = num_data - best_lc;
    out.right_sum_gradient = sum_gradient - best_slg;
    out.right_sum_hessian = sum_hessian - best_slh - K_EPSILON;
    out.gain = best_gain - min_gain_shift;
    out.cat_threshold.clear();
    if (use_onehot) out.cat_threshold.push_back((uint32_t)(best_threshold + offset));
    else if (best_dir == 1)
      for (int i = 0; i <= best_threshold; ++i) out.cat_threshold.push_back((uint32_t)(sorted_idx[i] + offset));
    else
      for (int i = 0; i <= best_threshold; ++i)
        out.cat_threshold.push_back((uint32_t)(sorted_idx[used_bin - 1 - i] + offset));
  }
  return is_splittable;
}

// split.rs:246-271
static bool find_best_threshold(const double* hist, const FeatureMeta& meta, const SplitParams& p, double sg,
                                double sh, int nd, double po, SplitInfo& out) {
  out.default_left = true;
  out.gain = K_MIN_SCORE;
  double shp = sh + 2.0 * K_EPSILON;
  bool s = meta.bin_type == BT_CAT ? find_best_threshold_categorical(hist, meta, p, sg, shp, nd, po, out)
                                   : find_best_threshold_numerical(hist, meta, p, sg, shp, nd, po, out);
  out.gain *= meta.penalty;
  return s;
}

// split.rs:893-925
static inline int32_t int_grad(int64_t x) { return (int32_t)(x >> 32); }
static inline uint32_t int_hess(int64_t x) { return (uint32_t)x; }
struct IntSums {
  int64_t sum;
  double grad_scale, hess_scale;
  double gradient() const { return (double)int_grad(sum) * grad_scale; }
  double hessian() const { return (double)int_hess(sum) * hess_scale; }
  double cnt_factor(int nd) const { return (double)nd / (double)int_hess(sum); }
};

static inline int64_t wadd(int64_t a, int64_t b) { return (int64_t)((uint64_t)a + (uint64_t)b); }
static inline int64_t wsub(int64_t a, int64_t b) { return (int64_t)((uint64_t)a - (uint64_t)b); }

// split.rs:993-1190
static bool scan_int(const int64_t* hist, const FeatureMeta& meta, const SplitParams& p, IntSums sums, int num_data,
                     double min_gain_shift, double parent_output, bool reverse, bool skip_default_bin,
                     bool na_as_missing, SplitInfo& out) {
  int offset = meta.offset;
  double gs = sums.grad_scale, hs = sums.hess_scale;
  int64_t local_sum = sums.sum;
  int64_t best_sum_left = 0;
  double best_gain = K_MIN_SCORE;
  uint32_t best_threshold = (uint32_t)meta.num_bin;
  double cnt_factor = sums.cnt_factor(num_data);
  int min_data = p.min_data_in_leaf;
  double min_hess = p.min_sum_hessian_in_leaf;
  bool is_splittable = false;
  if (reverse) {
    int64_t sum_right = 0;
    int t = meta.num_bin - 1 - offset - (int)na_as_missing;
    int t_end = 1 - offset;
    while (t >= t_end) {
      if (skip_default_bin && (t + offset) == (int)meta.default_bin) { t--; continue; }
      sum_right = wadd(sum_right, hist[t]);
      int rc = round_int((double)int_hess(sum_right) * cnt_factor);
      double srh = (double)int_hess(sum_right) * hs;
      if (rc < min_data || srh < min_hess) { t--; continue; }
      int lc = num_data - rc;
      if (lc < min_data) break;
      int64_t sum_left = wsub(local_sum, sum_right);
      double slh = (double)int_hess(sum_left) * hs;
      if (slh < min_hess) break;
      double srg = (double)int_grad(sum_right) * gs;
      double slg = (double)int_grad(sum_left) * gs;
      double cg = split_gain(slg, slh + K_EPSILON, srg, srh + K_EPSILON, p, lc, rc, parent_output);
      if (cg <= min_gain_shift) { t--; continue; }
      is_splittable = true;
      if (cg > best_gain) { best_sum_left = sum_left; best_threshold = (uint32_t)(t - 1 + offset); best_gain = cg; }
      t--;
    }
  } else {
    int64_t sum_left = 0;
    int t = 0;
    int t_end = meta.num_bin - 2 - offset;
    if (na_as_missing && offset == 1) {
      sum_left = local_sum;
      for (int i = 0; i < meta.num_bin - offset; ++i) sum_left = wsub(sum_left, hist[i]);
      t = -1;
    }
    while (t <= t_end) {
      if (skip_default_bin && (t + offset) == (int)meta.default_bin) { t++; continue; }
      if (t >= 0) sum_left = wadd(sum_left, hist[t]);
      int lc = round_int((double)int_hess(sum_left) * cnt_factor);
      double slh = (double)int_hess(sum_left) * hs;
      if (lc < min_data || slh < min_hess) { t++; continue; }
      int rc = num_data - lc;
      if (rc < min_data) break;
      int64_t sum_right = wsub(local_sum, sum_left);
      double srh = (double)int_hess(sum_right) * hs;
      if (srh < min_hess) break;
      double srg = (double)int_grad(sum_right) * gs;
      double slg = (double)int_grad(sum_left) * gs;
      double cg = split_gain(slg, slh + K_EPSILON, srg, srh + K_EPSILON, p, lc, rc, parent_output);
      if (cg <= min_gain_shift) { t++; continue; }
      is_splittable = true;
      if (cg > best_gain) { best_sum_left = sum_left; best_threshold = (uint32_t)(t + offset); best_gain = cg; }
      t++;
    }
  }
  if (is_splittable && best_gain > out.gain + min_gain_shift) {
    int64_t left = best_sum_left, right = wsub(sums.sum, left);
    double lg = (double)int_grad(left) * gs, lh = (double)int_hess(left) * hs;
    double rg = (double)int_grad(right) * gs, rh = (double)int_hess(right) * hs;
    int lc = round_int((double)int_hess(left) * cnt_factor);
    int rc = round_int((double)int_hess(right) * cnt_factor);
    out.threshold = best_threshold;
    out.left_output = leaf_output(lg, lh, p, lc, parent_output);
    out.left_count = lc;
    out.left_sum_gradient = lg;
    out.left_sum_hessian = lh;
    out.left_sum_gh = left;
    out.right_output = leaf_output(rg, rh, p, rc, parent_output);
    out.right_count = rc;
    out.right_sum_gradient = rg;
    out.right_sum_hessian = rh;
    out.right_sum_gh = right;
    out.gain = best_gain - min_gain_shift;
    out.default_left = reverse;
  }
  return is_splittable;
}

// split.rs:1194-1400 (categorical_int, no monotone constraints, no
// extra_trees): the categorical search on packed sums. Counts come from
// integer hessians; the chosen set's counts are recomputed from its packed
// sum (split.rs:1364-1367), not accumulated; no kEpsilon anywhere.
static bool categorical_int(const int64_t* hist, const FeatureMeta& meta, const SplitParams& p, IntSums sums,
                            int num_data, double parent_output, SplitInfo& out) {
  const double gs = sums.grad_scale, hs = sums.hess_scale;
  bool is_splittable = false;
  out.default_left = false;
  double best_gain = K_MIN_SCORE;
  int64_t best_sum_left = 0;
  const int64_t local_sum = sums.sum;
  const double sum_gradient = sums.gradient(), sum_hessian = sums.hessian();
  double gain_shift = p.use_smoothing() ? leaf_gain_given_output(sum_gradient, sum_hessian, p, parent_output)
                                        : leaf_gain(sum_gradient, sum_hessian, p, num_data, 0.0);
  double min_gain_shift = gain_shift + p.min_gain_to_split;
  int offset = meta.offset;
  int bin_start = 1 - offset, bin_end = meta.num_bin - offset;
  int used_bin = -1;
  std::vector<int> sorted_idx;
  SplitParams pc = p;
  bool use_onehot = meta.num_bin <= p.max_cat_to_onehot;
  int best_threshold = -1, best_dir = 1;
  const double cnt_factor = sums.cnt_factor(num_data);
  auto bin_cnt = [&](int64_t gh) { return round_int((double)int_hess(gh) * cnt_factor); };
  if (use_onehot) {
    for (int t = bin_start; t < bin_end; ++t) {
      int64_t gh = hist[t];
      int cnt = bin_cnt(gh);
      double hess = (double)int_hess(gh) * hs;
      if (cnt < p.min_data_in_leaf || hess < p.min_sum_hessian_in_leaf) continue;
      int other_count = num_data - cnt;
      if (other_count < p.min_data_in_leaf) continue;
      int64_t sum_other = wsub(local_sum, gh);
      double soh = (double)int_hess(sum_other) * hs;
      if (soh < p.min_sum_hessian_in_leaf) continue;
      double grad = (double)int_grad(gh) * gs;
      double sog = (double)int_grad(sum_other) * gs;
      double cg = split_gain(sog, soh, grad, hess, pc, other_count, cnt, parent_output);
      if (cg <= min_gain_shift) continue;
      is_splittable = true;
      if (cg > best_gain) { best_threshold = t; best_sum_left = gh; best_gain = cg; }
    }
  } else {
    for (int i = bin_start; i < bin_end; ++i)
      if ((double)bin_cnt(hist[i]) >= p.cat_smooth) sorted_idx.push_back(i);
    used_bin = (int)sorted_idx.size();
    pc.lambda_l2 += p.cat_l2;
    auto ctr = [&](int i) {
      return ((double)int_grad(hist[i]) * gs) / ((double)int_hess(hist[i]) * hs + p.cat_smooth);
    };
    std::stable_sort(sorted_idx.begin(), sorted_idx.end(), [&](int i, int j) { return ctr(i) < ctr(j); });
    int max_num_cat = std::min(p.max_cat_threshold, (used_bin + 1) / 2);
    is_splittable = false;
    const int dirs[2] = {1, -1};
    for (int d = 0; d < 2; ++d) {
      int dir = dirs[d];
      int start_pos = dir == 1 ? 0 : used_bin - 1;
      int cnt_cur_group = 0;
      int64_t sum_left = 0;
      int lc = 0;
      int i = 0;
      while (i < used_bin && i < max_num_cat) {
        int t = sorted_idx[start_pos];
        start_pos += dir;
        int64_t gh = hist[t];
        int cnt = bin_cnt(gh);
        sum_left = wadd(sum_left, gh);
        lc += cnt;
        cnt_cur_group += cnt;
        double slh = (double)int_hess(sum_left) * hs;
        if (lc < p.min_data_in_leaf || slh < p.min_sum_hessian_in_leaf) { i++; continue; }
        int rc = num_data - lc;
        if (rc < p.min_data_in_leaf || rc < p.min_data_per_group) break;
        int64_t sum_right = wsub(local_sum, sum_left);
        double srh = (double)int_hess(sum_right) * hs;
        if (srh < p.min_sum_hessian_in_leaf) break;
        if (cnt_cur_group < p.min_data_per_group) { i++; continue; }
        cnt_cur_group = 0;
        double slg = (double)int_grad(sum_left) * gs;
        double srg = (double)int_grad(sum_right) * gs;
        double cg = split_gain(slg, slh, srg, srh, pc, lc, rc, parent_output);
        if (cg <= min_gain_shift) { i++; continue; }
        is_splittable = true;
        if (cg > best_gain) { best_sum_left = sum_left; best_threshold = i; best_gain = cg; best_dir = dir; }
        i++;
      }
    }
  }
  if (is_splittable) {
    int64_t left = best_sum_left, right = wsub(local_sum, left);
    double lg = (double)int_grad(left) * gs, lh = (double)int_hess(left) * hs;
    double rg = (double)int_grad(right) * gs, rh = (double)int_hess(right) * hs;
    int lc = bin_cnt(left), rc = bin_cnt(right);
    out.left_output = leaf_output(lg, lh, pc, lc, parent_output);
    out.left_count = lc;
    out.left_sum_gradient = lg;
    out.left_sum_hessian = lh;
    out.right_output = leaf_output(rg, rh, pc, rc, parent_output);
    out.right_count = rc;
    out.right_sum_gradient = rg;
    out.right_sum_hessian = rh;
    out.gain = best_gain - min_gain_shift;
    out.left_sum_gh = left;
    out.right_sum_gh = right;
    out.cat_threshold.clear();
    if (use_onehot) out.cat_threshold.push_back((uint32_t)(best_threshold + offset));
    else if (best_dir == 1)
      for (int i = 0; i <= best_threshold; ++i) out.cat_threshold.push_back((uint32_t)(sorted_idx[i] + offset));
    else
      for (int i = 0; i <= best_threshold; ++i)
        out.cat_threshold.push_back((uint32_t)(sorted_idx[used_bin - 1 - i] + offset));
  }
  return is_splittable;
}

// split.rs:928-990
static bool find_best_threshold_int(const int64_t* hist, const FeatureMeta& meta, const SplitParams& p, IntSums sums,
                                    int nd, double po, SplitInfo& out) {
  out.default_left = true;
  out.gain = K_MIN_SCORE;
  if (meta.bin_type == BT_CAT) {
    bool s = categorical_int(hist, meta, p, sums, nd, po, out);
    out.gain *= meta.penalty;
    return s;
  }
  double min_gain_shift = leaf_gain(sums.gradient(), sums.hessian(), p, nd, po) + p.min_gain_to_split;
  bool splittable = false;
  if (meta.num_bin > 2 && meta.missing_type != MT_NONE) {
    if (meta.missing_type == MT_ZERO) {
      splittable |= scan_int(hist, meta, p, sums, nd, min_gain_shift, po, true, true, false, out);
      splittable |= scan_int(hist, meta, p, sums, nd, min_gain_shift, po, false, true, false, out);
    } else {
      splittable |= scan_int(hist, meta, p, sums, nd, min_gain_shift, po, true, false, true, out);
      splittable |= scan_int(hist, meta, p, sums, nd, min_gain_shift, po, false, false, true, out);
    }
  } else {
    splittable |= scan_int(hist, meta, p, sums, nd, min_gain_shift, po, true, false, false, out);
    if (meta.missing_type == MT_NAN) out.default_left = false;
  }
  out.gain *= meta.penalty;
  return splittable;
}

// ---------------------------------------------------------------- histogram helpers
// histogram.rs:42-56; `hist` is the feature view ([g,h] per stored bin)
static void hist_fix(const FeatureMeta& m, double sg, double sh, double* hist) {
  uint32_t mfb = m.most_freq_bin;
  if (mfb > 0) {
    double g = sg, h = sh;
    for (int i = 0; i < m.num_bin; ++i) {
      if ((uint32_t)i != mfb) { g -= hist[2 * i]; h -= hist[2 * i + 1]; }
    }
    hist[2 * mfb] = g;
    hist[2 * mfb + 1] = h;
  }
}
// histogram.rs:67-78
static void hist_fix_int(const FeatureMeta& m, int64_t s, int64_t* hist) {
  uint32_t mfb = m.most_freq_bin;
  if (mfb > 0) {
    for (int i = 0; i < m.num_bin; ++i) if ((uint32_t)i != mfb) s = wsub(s, hist[i]);
    hist[mfb] = s;
  }
}

// ---------------------------------------------------------------- tree
// tree.rs:120-159, 254-334
struct Tree {
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

  static double round0(double v) { return (v >= -K_ZERO_THRESHOLD && v <= K_ZERO_THRESHOLD) ? 0.0 : v; }
  int split(int leaf, int fi, int fr, uint32_t tbin, double thr, double lv, double rv, int lc, int rc, double lw,
            double rw, float gain, int mt, bool dl) {
    int new_node = num_leaves - 1, new_leaf = num_leaves;
    int parent = leaf_parent[leaf];
    if (parent >= 0) {
      if (left_child[parent] == ~leaf) left_child[parent] = new_node; else right_child[parent] = new_node;
    }
    split_feature_inner.push_back(fi);
    split_feature.push_back(fr);
    split_gain.push_back(gain);
    left_child.push_back(~leaf);
    right_child.push_back(~new_leaf);
    leaf_parent[leaf] = new_node;
    leaf_parent.push_back(new_node);
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
    int8_t dt = 0;
    if (dl) dt |= 2;
    dt = (int8_t)((dt & 3) | (mt << 2));
    decision_type.push_back(dt);
    threshold_in_bin.push_back(tbin);
    threshold.push_back(thr);
    num_leaves += 1;
    return new_leaf;
  }
  void set_leaf_output(int leaf, double v) { leaf_value[leaf] = round0(v); }
  void shrink(double rate) {
    for (int i = 0; i < num_leaves - 1; ++i) {
      leaf_value[i] = round0(leaf_value[i] * rate);
      internal_value[i] = round0(internal_value[i] * rate);
    }
    leaf_value[num_leaves - 1] = round0(leaf_value[num_leaves - 1] * rate);
    shrinkage *= rate;
  }
  void add_bias(double v) {
    for (int i = 0; i < num_leaves - 1; ++i) {
      leaf_value[i] = round0(leaf_value[i] + v);
      internal_value[i] = round0(internal_value[i] + v);
    }
    leaf_value[num_leaves - 1] = round0(leaf_value[num_leaves - 1] + v);
    shrinkage = 1.0;
  }
  void as_constant(double v, int count) {
    num_leaves = 1;
    shrinkage = 1.0;
    leaf_value.assign(1, v);
    leaf_count.assign(1, count);
    leaf_weight.resize(1);
    leaf_parent.resize(1);
    leaf_depth.resize(1);
    split_feature_inner.clear(); split_feature.clear(); left_child.clear(); right_child.clear();
    internal_count.clear(); split_gain.clear(); threshold_in_bin.clear(); threshold.clear();
    decision_type.clear(); internal_value.clear(); internal_weight.clear();
  }
  // tree.rs:381-392, 417-443
  int numerical_decision(double fval, int node) const {
    int mt = (decision_type[node] >> 2) & 3;
    bool dl = (decision_type[node] & 2) != 0;
    if (std::isnan(fval) && mt != MT_NAN) fval = 0.0;
    bool zero = fval >= -K_ZERO_THRESHOLD && fval <= K_ZERO_THRESHOLD;
    if ((mt == MT_ZERO && zero) || (mt == MT_NAN && std::isnan(fval))) return dl ? left_child[node] : right_child[node];
    return fval <= threshold[node] ? left_child[node] : right_child[node];
  }
  int categorical_decision(double fval, int node) const {
    if (std::isnan(fval)) return right_child[node];
    int v = (int)fval;
    int c = (int)threshold[node];
    if (v >= 0) {
      int i1 = v / 32;
      int lo = cat_boundaries[c], hi = cat_boundaries[c + 1];
      if (i1 < hi - lo && ((cat_threshold[lo + i1] >> (v % 32)) & 1)) return left_child[node];
    }
    return right_child[node];
  }
  int get_leaf(const double* row, int ncol) const {
    if (num_leaves <= 1) return 0;
    int node = 0;
    while (node >= 0) {
      int f = split_feature[node];
      double v = f < ncol ? row[f] : 0.0;
      node = (decision_type[node] & 1) ? categorical_decision(v, node) : numerical_decision(v, node);
    }
    return ~node;
  }
};

// tree.rs:27-44 ConstructBitset / FindInBitset
static std::vector<uint32_t> construct_bitset(const std::vector<int>& vals) {
  std::vector<uint32_t> ret;
  for (int v : vals) {
    size_t i1 = (size_t)(v / 32);
    if (ret.size() < i1 + 1) ret.resize(i1 + 1, 0);
    ret[i1] |= 1u << (v % 32);
  }
  return ret;
}
static bool find_in_bitset(const std::vector<uint32_t>& bits, int pos) {
  size_t i1 = (size_t)(pos / 32);
  if (i1 >= bits.size()) return false;
  return ((bits[i1] >> (pos % 32)) & 1) == 1;
}

// Host threads of the CPU engine (orc_set_num_threads; 1 by default). Every
// parallel loop below splits work whose results do not depend on the split:
// histograms by feature (each bin still sums its rows in partition order),
// threshold searches by feature (the best split is then picked in feature
// order), the stable partition by row chunks, per-row gradients and
// quantization, per-leaf score updates and renewals. So trees and scores are
// bit-identical at every thread count.
static int g_threads = 1;

// ---------------------------------------------------------------- quantizer
// quant.rs:37-40
static inline int8_t to_i8(double x) {
  int32_t t = (std::isnan(x) || !(x >= -2147483648.0 && x < 2147483648.0)) ? INT32_MIN : (int32_t)x;
  return (int8_t)t;
}

// quant.rs:16-159 (num_threads = 1)
struct Discretizer {
  int bins;
  bool constant_hessian = false, stochastic;
  std::vector<double> grv, hrv;
  Mt19937 start_eng;
  int start_max;
  std::vector<int8_t> gq, hq;
  std::vector<int64_t> packed;
  double grad_scale = 0, hess_scale = 0;
  Discretizer(int b, int seed, bool stoch, int64_t n, int num_threads)
      : bins(b), stochastic(stoch), start_eng((uint32_t)seed), start_max((int)n) {
    grv.assign(n, 0.0);
    hrv.assign(n, 0.0);
    // multi_val_bin.rs:337-345 block_info(num_threads, n, 512)
    int64_t nblock = std::max<int64_t>(1, std::min<int64_t>(num_threads, (n + 511) / 512));
    int64_t bsize = n;
    if (nblock > 1) { bsize = (n + nblock - 1) / nblock; bsize = (bsize + 31) / 32 * 32; }
    for (int64_t b = 0; b < nblock; ++b) {
      int64_t s = b * bsize, e = std::min(n, s + bsize);
      Mt19937 ge((uint32_t)(seed + (int)b));
      Mt19937 he((uint32_t)(seed + (int)b + num_threads));
      for (int64_t i = s; i < e; ++i) { grv[i] = ge.uniform01(); hrv[i] = he.uniform01(); }
    }
    gq.assign(n, 0);
    hq.assign(n, 0);
    packed.assign(n, 0);
  }
  // quant.rs:92-129 with rows: position i of the bag is row rows[i]; the
  // random tables are read at (i + start) mod (bag size).
  void discretize_rows(const float* grad, const float* hess, const std::vector<uint32_t>& rows) {
    const int64_t n = (int64_t)rows.size();
    if (n == 0) return;
    double mg = std::fabs((double)grad[rows[0]]), mh = std::fabs((double)hess[rows[0]]);
    for (int64_t i = 0; i < n; ++i) {
      double g = std::fabs((double)grad[rows[i]]), h = std::fabs((double)hess[rows[i]]);
      if (g > mg) mg = g;
      if (h > mh) mh = h;
    }
    grad_scale = mg / (double)(bins / 2);
    hess_scale = constant_hessian ? mh : mh / (double)bins;
    double inv_g = 1.0 / grad_scale, inv_h = 1.0 / hess_scale;
    int64_t start = start_eng.uniform_int(0, start_max);
    for (int64_t i = 0; i < n; ++i) {
      const uint32_t r = rows[i];
      double g = (double)grad[r];
      double rg = 0.5, rh = 0.5;
      if (stochastic) { int64_t pos = (i + start) % n; rg = grv[pos]; rh = hrv[pos]; }
      int8_t gi = g >= 0.0 ? to_i8(g * inv_g + rg) : to_i8(g * inv_g - rg);
      int8_t hi = constant_hessian ? 1 : to_i8((double)hess[r] * inv_h + rh);
      gq[r] = gi;
      hq[r] = hi;
      packed[r] = (int64_t)(((uint64_t)(int64_t)gi << 32) | (uint64_t)(uint8_t)hi);
    }
  }
  void discretize(const float* grad, const float* hess, int64_t n) {
    if (n == 0) return;
    double mg = std::fabs((double)grad[0]), mh = std::fabs((double)hess[0]);
    // a maximum does not depend on the order of its operands
#pragma omp parallel for num_threads(g_threads) schedule(static) reduction(max : mg, mh) if (g_threads > 1)
    for (int64_t i = 0; i < n; ++i) {
      double g = std::fabs((double)grad[i]), h = std::fabs((double)hess[i]);
      if (g > mg) mg = g;
      if (h > mh) mh = h;
    }
    grad_scale = mg / (double)(bins / 2);
    hess_scale = constant_hessian ? mh : mh / (double)bins;
    double inv_g = 1.0 / grad_scale, inv_h = 1.0 / hess_scale;
    int64_t start = start_eng.uniform_int(0, start_max);
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
    for (int64_t i = 0; i < n; ++i) {
      double g = (double)grad[i];
      double rg = 0.5, rh = 0.5;
      if (stochastic) { int64_t pos = (i + start) % n; rg = grv[pos]; rh = hrv[pos]; }
      int8_t gi = g >= 0.0 ? to_i8(g * inv_g + rg) : to_i8(g * inv_g - rg);
      int8_t hi = constant_hessian ? 1 : to_i8((double)hess[i] * inv_h + rh);
      gq[i] = gi;
      hq[i] = hi;
      packed[i] = (int64_t)(((uint64_t)(int64_t)gi << 32) | (uint64_t)(uint8_t)hi);
    }
  }
};

// ---------------------------------------------------------------- learner
struct LeafSplits {
  int leaf = -1, num_data = 0;
  double sum_gradients = 0, sum_hessians = 0;
  int64_t int_sum = 0;
  double weight = 0;
};

struct TraceEntry { int leaf, feature; uint32_t threshold; double gain; };

// learner/mod.rs:86-153, 572-706, 816-1225 (single thread, no bagging,
// feature_fraction = 1, no constraints/CEGB/forced splits/linear trees)

struct Learner {
  const Dataset* ds;
  Params prm;
  SplitParams sp;
  int num_leaves;
  int F, B;
  std::vector<FeatureMeta> metas;
  // partition (partition.rs)
  std::vector<uint32_t> indices;
  std::vector<int64_t> leaf_begin, leaf_count;
  // per-leaf histogram pool (bin-indexed [F][B][2] / [F][B])
  std::vector<std::vector<double>> pool;
  std::vector<std::vector<int64_t>> ipool;
  std::vector<std::vector<char>> splittable;
  std::vector<SplitInfo> best;
  LeafSplits smaller, larger;
  Discretizer* quant = nullptr;
  std::vector<TraceEntry> trace;
  LeafSplits root_sums;   // the last tree's root sums (stage checks)
  // learner/mod.rs:545-568 set_bagging_data: the root's rows (ascending);
  // `bag_subset` is upstream's subset mode (the quantizer draws over the bag)
  std::vector<uint32_t> bag;
  bool use_bag = false, bag_subset = false;
  // the last tree's smaller-leaf histograms, for stage parity checks
  Learner(const Dataset* d, const Params& p) : ds(d), prm(p) {
    sp = SplitParams{p.lambda_l1, p.lambda_l2, p.max_delta_step, p.path_smooth, p.min_data_in_leaf,
                     p.min_sum_hessian_in_leaf, p.min_gain_to_split, p.max_cat_to_onehot, p.max_cat_threshold,
                     p.cat_l2, p.cat_smooth, p.min_data_per_group};
    num_leaves = std::max(p.num_leaves, 2);
    F = (int)d->used.size();
    B = 1;
    for (int f = 0; f < F; ++f) {
      const BinMapper& m = d->mappers[d->used[f]];
      FeatureMeta fm;
      fm.num_bin = m.num_bin;
      fm.missing_type = m.missing_type;
      fm.offset = m.most_freq_bin == 0 ? 1 : 0;
      fm.default_bin = m.default_bin;
      fm.most_freq_bin = m.most_freq_bin;
      fm.bin_type = m.bin_type;
      metas.push_back(fm);
      B = std::max(B, m.num_bin);
    }
    indices.resize(d->n);
    leaf_begin.assign(num_leaves, 0);
    leaf_count.assign(num_leaves, 0);
    pool.resize(num_leaves);
    ipool.resize(num_leaves);
    splittable.assign(num_leaves, std::vector<char>(F, 1));
    best.resize(num_leaves);
    if (p.use_quantized_grad) quant = new Discretizer(p.num_grad_quant_bins, p.seed, p.stochastic_rounding, d->n, p.num_threads);
  }
  ~Learner() { delete quant; }

  // The feature view of leaf buffer `buf` (bin-indexed storage, view
  // index t = bin - offset).
  double* view(std::vector<double>& buf, int f) { return buf.data() + (size_t)f * B * 2 + 2 * metas[f].offset; }
  int64_t* iview(std::vector<int64_t>& buf, int f) { return buf.data() + (size_t)f * B + metas[f].offset; }

  // multi_val_bin.rs:623-656 with one block: rows in partition order,
  // every bin summed sequentially; the most frequent bin is not stored.
  void construct_f64(int leaf, const float* g, const float* h, std::vector<double>& out) {
    out.assign((size_t)F * B * 2, 0.0);
    int64_t b = leaf_begin[leaf], c = leaf_count[leaf];
    if (g_threads > 1 && c >= 4096) {
      // by feature range: one pass over the rows per thread, and every bin of
      // its features sums the rows in the same order as the loop below
      const int T = std::min(g_threads, F);
#pragma omp parallel num_threads(T)
      {
        const int t = omp_get_thread_num(), f0 = F * t / T, f1 = F * (t + 1) / T;
        for (int64_t k = 0; k < c; ++k) {
          uint32_t r = indices[b + k];
          double gg = (double)g[r], hh = (double)h[r];
          for (int f = f0; f < f1; ++f) {
            uint32_t bin = ds->bin(r, f);
            if (bin == metas[f].most_freq_bin) continue;
            size_t o = ((size_t)f * B + bin) * 2;
            out[o] += gg;
            out[o + 1] += hh;
          }
        }
      }
      return;
    }
    for (int64_t k = 0; k < c; ++k) {
      uint32_t r = indices[b + k];
      double gg = (double)g[r], hh = (double)h[r];
      for (int f = 0; f < F; ++f) {
        uint32_t bin = ds->bin(r, f);
        if (bin == metas[f].most_freq_bin) continue;
        size_t o = ((size_t)f * B + bin) * 2;
        out[o] += gg;
        out[o + 1] += hh;
      }
    }
  }
  // multi_val_bin.rs:763-791
  void construct_int(int leaf, std::vector<int64_t>& out) {
    out.assign((size_t)F * B, 0);
    int64_t b = leaf_begin[leaf], c = leaf_count[leaf];
    if (g_threads > 1 && c >= 4096) {
      // by row chunks into private histograms, then summed: wrapping integer
      // sums do not depend on the order of their terms
      const int T = g_threads;
      std::vector<std::vector<int64_t>> part(T);
#pragma omp parallel num_threads(T)
      {
        const int t = omp_get_thread_num();
        std::vector<int64_t>& o = part[t];
        o.assign((size_t)F * B, 0);
        for (int64_t k = c * t / T; k < c * (t + 1) / T; ++k) {
          uint32_t r = indices[b + k];
          int64_t v = quant->packed[r];
          for (int f = 0; f < F; ++f) {
            uint32_t bin = ds->bin(r, f);
            if (bin == metas[f].most_freq_bin) continue;
            o[(size_t)f * B + bin] = wadd(o[(size_t)f * B + bin], v);
          }
        }
      }
      for (int t = 0; t < T; ++t)
        for (size_t i = 0; i < out.size(); ++i) out[i] = wadd(out[i], part[t][i]);
      return;
    }
    for (int64_t k = 0; k < c; ++k) {
      uint32_t r = indices[b + k];
      int64_t v = quant->packed[r];
      for (int f = 0; f < F; ++f) {
        uint32_t bin = ds->bin(r, f);
        if (bin == metas[f].most_freq_bin) continue;
        out[(size_t)f * B + bin] = wadd(out[(size_t)f * B + bin], v);
      }
    }
  }

  double parent_output(const Tree& t, const LeafSplits& ls) {
    return t.num_leaves == 1 ? root_output(ls.sum_gradients, ls.sum_hessians, sp, ls.num_data) : ls.weight;
  }

  // learner/mod.rs:816-1103
  void find_best_splits(const Tree& tree, const float* g, const float* h, int left, int right) {
    bool q = quant != nullptr;
    int sm = smaller.leaf;
    std::vector<char> parent_spl;
    if (right >= 0) parent_spl = splittable[left];
    std::vector<double> sbuf;
    std::vector<int64_t> sibuf;
    if (q) construct_int(sm, sibuf); else construct_f64(sm, g, h, sbuf);
    std::vector<double> lbuf;
    std::vector<int64_t> libuf;
    if (right >= 0) { if (q) libuf = ipool[left]; else lbuf = pool[left]; }
    double spo = parent_output(tree, smaller);
    double lpo = larger.leaf >= 0 ? parent_output(tree, larger) : 0.0;
    SplitInfo s_best, l_best;
    std::vector<char> s_spl(F, 0), l_spl = right >= 0 ? parent_spl : std::vector<char>(F, 0);
    std::vector<SplitInfo> s_res(F), l_res(F);
    std::vector<char> s_has(F, 0);
    // each feature's search is independent; the best is picked below in feature order
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 1) if (g_threads > 1)
    for (int f = 0; f < F; ++f) {
      bool used = right >= 0 ? parent_spl[f] != 0 : true;
      if (!used) { s_spl[f] = 0; continue; }
      const FeatureMeta& m = metas[f];
      int real = ds->used[f];
      SplitInfo ss, ls;
      bool s_ok, l_ok = false;
      if (q) {
        int64_t* hv = iview(sibuf, f);
        int64_t* hb = sibuf.data() + (size_t)f * B;
        hist_fix_int(m, smaller.int_sum, hb);
        s_ok = find_best_threshold_int(hv, m, sp, IntSums{smaller.int_sum, quant->grad_scale, quant->hess_scale},
                                       smaller.num_data, spo, ss);
        ss.feature = real;
        if (larger.leaf >= 0) {
          int64_t* lb = libuf.data() + (size_t)f * B;
          for (int i = 0; i < B; ++i) lb[i] = wsub(lb[i], hb[i]);
          l_ok = find_best_threshold_int(iview(libuf, f), m, sp,
                                         IntSums{larger.int_sum, quant->grad_scale, quant->hess_scale},
                                         larger.num_data, lpo, ls);
          ls.feature = real;
        }
      } else {
        double* hb = sbuf.data() + (size_t)f * B * 2;
        hist_fix(m, smaller.sum_gradients, smaller.sum_hessians, hb);
        s_ok = find_best_threshold(view(sbuf, f), m, sp, smaller.sum_gradients, smaller.sum_hessians,
                                   smaller.num_data, spo, ss);
        ss.feature = real;
        if (larger.leaf >= 0) {
          double* lb = lbuf.data() + (size_t)f * B * 2;
          for (int i = 0; i < 2 * B; ++i) lb[i] -= hb[i];
          l_ok = find_best_threshold(view(lbuf, f), m, sp, larger.sum_gradients, larger.sum_hessians,
                                     larger.num_data, lpo, ls);
          ls.feature = real;
        }
      }
      s_res[f] = ss;
      l_res[f] = ls;
      s_has[f] = 1;
      s_spl[f] = s_ok;
      if (larger.leaf >= 0) l_spl[f] = l_ok;
    }
    for (int f = 0; f < F; ++f) {
      if (!s_has[f]) continue;
      if (s_res[f].better_than(s_best)) s_best = s_res[f];
      if (larger.leaf >= 0 && l_res[f].better_than(l_best)) l_best = l_res[f];
    }
    best[sm] = s_best;
    if (q) ipool[sm] = sibuf; else pool[sm] = sbuf;
    splittable[sm] = s_spl;
    if (larger.leaf >= 0) {
      best[larger.leaf] = l_best;
      if (q) ipool[larger.leaf] = libuf; else pool[larger.leaf] = lbuf;
      splittable[larger.leaf] = l_spl;
    }
  }

  bool before_find_best_split(const Tree& t, int left, int right) {
    if (prm.max_depth > 0 && t.leaf_depth[left] >= prm.max_depth) {
      best[left].gain = K_MIN_SCORE;
      if (right >= 0) best[right].gain = K_MIN_SCORE;
      return false;
    }
    int64_t nl = left >= 0 ? leaf_count[left] : 0, nr = right >= 0 ? leaf_count[right] : 0;
    int64_t min2 = (int64_t)prm.min_data_in_leaf * 2;
    if (nr < min2 && nl < min2) {
      best[left].gain = K_MIN_SCORE;
      if (right >= 0) best[right].gain = K_MIN_SCORE;
      return false;
    }
    return true;
  }

  // learner/mod.rs:792-810
  bool goes_left(uint32_t b, const FeatureMeta& m, uint32_t thr, bool dl) {
    bool miss = (m.missing_type == MT_ZERO && b == m.default_bin) || (m.missing_type == MT_NAN && b == (uint32_t)(m.num_bin - 1));
    return miss ? dl : b <= thr;
  }

  // partition.rs:107-174 (one thread)
  int64_t partition_split(int leaf, int f, const std::vector<char>& lut, int right_leaf) {
    int64_t b = leaf_begin[leaf], c = leaf_count[leaf];
    int64_t nl = 0;
    if (g_threads > 1 && c >= 8192) {
      // row chunks in order: each chunk's left / right rows go after the
      // earlier chunks' (the same stable partition as one pass)
      const int T = g_threads;
      std::vector<int64_t> lc(T + 1, 0);
      std::unique_ptr<uint32_t[]> tmp(new uint32_t[(size_t)c]);
#pragma omp parallel num_threads(T)
      {
        const int t = omp_get_thread_num();
        const int64_t lo = c * t / T, hi = c * (t + 1) / T;
        int64_t k = 0;
        for (int64_t i = lo; i < hi; ++i) k += lut[ds->bin(indices[b + i], f)] ? 1 : 0;
        lc[t + 1] = k;
#pragma omp barrier
#pragma omp single
        for (int u = 0; u < T; ++u) lc[u + 1] += lc[u];
        int64_t lp = lc[t], rp = lc[T] + (lo - lc[t]);
        for (int64_t i = lo; i < hi; ++i) {
          const uint32_t r = indices[b + i];
          if (lut[ds->bin(r, f)]) tmp[lp++] = r; else tmp[rp++] = r;
        }
      }
      std::memcpy(indices.data() + b, tmp.get(), (size_t)c * 4);
      nl = lc[T];
    } else {
      std::vector<uint32_t> L, R;
      for (int64_t k = 0; k < c; ++k) {
        uint32_t r = indices[b + k];
        if (lut[ds->bin(r, f)]) L.push_back(r); else R.push_back(r);
      }
      std::copy(L.begin(), L.end(), indices.begin() + b);
      std::copy(R.begin(), R.end(), indices.begin() + b + (int64_t)L.size());
      nl = (int64_t)L.size();
    }
    leaf_count[leaf] = nl;
    leaf_begin[right_leaf] = b + nl;
    leaf_count[right_leaf] = c - nl;
    return nl;
  }

  // learner/mod.rs:1105-1225
  std::pair<int, int> split(Tree& tree, int best_leaf) {
    SplitInfo info = best[best_leaf];
    int inner = ds->real_to_inner[info.feature];
    const BinMapper& mapper = ds->mappers[info.feature];
    int next_leaf = tree.num_leaves;
    const FeatureMeta& m = metas[inner];
    int nb = std::max(m.num_bin, 1);
    bool numerical = m.bin_type == BT_NUM;
    std::vector<char> lut(nb);
    std::vector<uint32_t> bits_inner, bits;
    if (numerical) {
      for (int b = 0; b < nb; ++b) lut[b] = goes_left((uint32_t)b, m, info.threshold, info.default_left);
    } else {
      std::vector<int> vi, vr;
      for (uint32_t b : info.cat_threshold) { vi.push_back((int)b); vr.push_back((int)mapper.bin_to_value(b)); }
      bits_inner = construct_bitset(vi);
      bits = construct_bitset(vr);
      for (int b = 0; b < nb; ++b) {
        if ((uint32_t)b == m.most_freq_bin) lut[b] = m.most_freq_bin > 0 && find_in_bitset(bits_inner, (int)m.most_freq_bin);
        else lut[b] = find_in_bitset(bits_inner, b);
      }
    }
    partition_split(best_leaf, inner, lut, next_leaf);
    info.left_count = (int)leaf_count[best_leaf];
    info.right_count = (int)leaf_count[next_leaf];
    float gain = (float)(info.gain + sp.min_gain_to_split);
    int right;
    if (numerical) {
      right = tree.split(best_leaf, inner, info.feature, info.threshold, mapper.bin_to_value(info.threshold),
                         info.left_output, info.right_output, info.left_count, info.right_count,
                         info.left_sum_hessian, info.right_sum_hessian, gain, mapper.missing_type, info.default_left);
    } else {
      right = tree.split(best_leaf, inner, info.feature, (uint32_t)tree.num_cat, (double)tree.num_cat,
                         info.left_output, info.right_output, info.left_count, info.right_count,
                         info.left_sum_hessian, info.right_sum_hessian, gain, mapper.missing_type, false);
      size_t node = tree.decision_type.size() - 1;
      tree.decision_type[node] |= 1;
      tree.num_cat += 1;
      tree.cat_boundaries.push_back(tree.cat_boundaries.back() + (int)bits.size());
      tree.cat_threshold.insert(tree.cat_threshold.end(), bits.begin(), bits.end());
      tree.cat_boundaries_inner.push_back(tree.cat_boundaries_inner.back() + (int)bits_inner.size());
      tree.cat_threshold_inner.insert(tree.cat_threshold_inner.end(), bits_inner.begin(), bits_inner.end());
    }
    trace.push_back({best_leaf, info.feature, info.threshold, info.gain});
    LeafSplits ls{best_leaf, info.left_count, info.left_sum_gradient, info.left_sum_hessian, info.left_sum_gh, info.left_output};
    LeafSplits rs{right, info.right_count, info.right_sum_gradient, info.right_sum_hessian, info.right_sum_gh, info.right_output};
    if (info.left_count < info.right_count) { smaller = ls; larger = rs; }
    else { smaller = rs; larger = ls; }
    return {best_leaf, right};
  }

  // learner/mod.rs:572-695
  Tree train(const float* g, const float* h) {
    int64_t n = use_bag ? (int64_t)bag.size() : ds->n;
    bool q = quant != nullptr;
    // a full-data bag (not subset mode) keeps the learner's view of every
    // row: the quantizer runs over all rows (quant.rs:92-129 with rows None)
    const bool full_bag = use_bag && !bag_subset;
    if (q) {
      if (full_bag) quant->discretize(g, h, ds->n);
      else if (use_bag) quant->discretize_rows(g, h, bag);
      else quant->discretize(g, h, n);
    }
    for (int64_t i = 0; i < n; ++i) indices[i] = use_bag ? bag[i] : (uint32_t)i;
    std::fill(leaf_begin.begin(), leaf_begin.end(), 0);
    std::fill(leaf_count.begin(), leaf_count.end(), 0);
    leaf_count[0] = n;
    for (auto& s : best) s.reset();
    double sg = 0, sh = 0;
    int64_t is = 0;
    if (q && full_bag) {
      // quant.rs:144-158 bag_root_sums: f32 scales, and the packed sum read
      // by position 0 .. n - 1 rather than over the bag's rows
      const float gs = (float)quant->grad_scale, hs = (float)quant->hess_scale;
      for (int64_t k = 0; k < n; ++k) {
        const uint32_t i = indices[k];
        sg += (double)((float)quant->gq[i] * gs);
        sh += (double)((float)quant->hq[i] * hs);
        is = wadd(is, quant->packed[k]);
      }
    } else if (q) {
      for (int64_t k = 0; k < n; ++k) {
        const uint32_t i = indices[k];
        sg += (double)quant->gq[i] * quant->grad_scale;
        sh += (double)quant->hq[i] * quant->hess_scale;
        is = wadd(is, quant->packed[i]);
      }
    } else {
      for (int64_t k = 0; k < n; ++k) { sg += (double)g[indices[k]]; sh += (double)h[indices[k]]; }
    }
    smaller = LeafSplits{0, (int)n, sg, sh, is, smaller.weight};
    root_sums = smaller;
    larger = LeafSplits{};
    trace.clear();
    Tree tree;
    tree.set_leaf_output(0, root_output(sg, sh, sp, (int)n));
    int left = 0, right = -1;
    for (int s = 0; s < num_leaves - 1; ++s) {
      if (before_find_best_split(tree, left, right)) find_best_splits(tree, g, h, left, right);
      int bl = 0;
      for (int i = 1; i < num_leaves; ++i) if (best[i].better_than(best[bl])) bl = i;
      double gain = best[bl].gain;
      if (!(gain > 0.0)) break;
      auto lr = split(tree, bl);
      left = lr.first;
      right = lr.second;
    }
    if (q && prm.quant_train_renew_leaf) {
      // learner/mod.rs:736-747
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 1) if (g_threads > 1)
      for (int leaf = 0; leaf < tree.num_leaves; ++leaf) {
        double a = 0, b = 0;
        for (int64_t k = 0; k < leaf_count[leaf]; ++k) {
          uint32_t r = indices[leaf_begin[leaf] + k];
          a += (double)g[r];
          b += (double)h[r];
        }
        tree.set_leaf_output(leaf, refit_leaf_output(a, b, sp, false, (int)leaf_count[leaf], 0.0));
      }
    }
    return tree;
  }

  // tree.rs:605-632 leaf_binned_with: the leaf of row r from its bins
  int leaf_binned(const Tree& t, int64_t r) const {
    if (t.num_leaves <= 1) return 0;
    int node = 0;
    while (node >= 0) {
      const int f = t.split_feature_inner[node];
      const FeatureMeta& m = metas[f];
      const uint32_t b = ds->bin(r, f);
      if (t.decision_type[node] & 1) {
        const int c = (int)t.threshold_in_bin[node];
        const int lo = t.cat_boundaries_inner[c], hi = t.cat_boundaries_inner[c + 1];
        const bool in = (int)(b / 32) < hi - lo && ((t.cat_threshold_inner[lo + b / 32] >> (b % 32)) & 1);
        node = in ? t.left_child[node] : t.right_child[node];
        continue;
      }
      const int mt = (t.decision_type[node] >> 2) & 3;
      const bool dl = (t.decision_type[node] & 2) != 0;
      const bool def = (mt == MT_ZERO && b == (uint32_t)m.default_bin) || (mt == MT_NAN && b == (uint32_t)(m.num_bin - 1));
      node = def ? (dl ? t.left_child[node] : t.right_child[node])
                 : (b <= t.threshold_in_bin[node] ? t.left_child[node] : t.right_child[node]);
    }
    return ~node;
  }

  // partition.rs:178-191
  void add_leaf_outputs(const Tree& t, double* score) {
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 1) if (g_threads > 1)
    for (int leaf = 0; leaf < t.num_leaves; ++leaf)
      for (int64_t k = 0; k < leaf_count[leaf]; ++k) score[indices[leaf_begin[leaf] + k]] += t.leaf_value[leaf];
  }
};

// ---------------------------------------------------------------- objectives
// objective/binary.rs:101-127
static void binary_gradients(const double* s, const float* label, int64_t n, double sigmoid, float* g, float* h) {
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
  for (int64_t i = 0; i < n; ++i) {
    bool pos = label[i] > 0.0f;
    double lab = pos ? 1.0 : -1.0;
    double response = -lab * sigmoid / ((double)1.0f + std::exp(lab * sigmoid * s[i]));
    double ar = std::fabs(response);
    g[i] = (float)(response * 1.0);
    h[i] = (float)(ar * (sigmoid - ar) * 1.0);
  }
}
// objective/binary.rs:129-154
static double binary_boost_from_score(const float* label, int64_t n, double sigmoid) {
  double suml = 0.0, sumw = (double)n;
  for (int64_t i = 0; i < n; ++i) suml += (double)(int)(label[i] > 0.0f);
  double pavg = suml / sumw;
  pavg = std::min(pavg, 1.0 - K_EPSILON);
  pavg = std::max(pavg, K_EPSILON);
  return std::log(pavg / ((double)1.0f - pavg)) / sigmoid;
}
// objective/multiclass.rs:16-30, 126-154
static void multiclass_gradients(const double* s, const float* label, int64_t n, int K, float* g, float* h) {
  double factor = (double)K / (double)((float)K - 1.0f);
  std::vector<double> raw(K), rec(K);
  for (int64_t i = 0; i < n; ++i) {
    for (int c = 0; c < K; ++c) raw[c] = s[(size_t)c * n + i];
    double wmax = raw[0];
    for (int c = 1; c < K; ++c) wmax = raw[c] < wmax ? wmax : raw[c];
    double wsum = 0.0;
    for (int c = 0; c < K; ++c) { rec[c] = std::exp(raw[c] - wmax); wsum += rec[c]; }
    for (int c = 0; c < K; ++c) rec[c] /= wsum;
    int li = (int)label[i];
    for (int c = 0; c < K; ++c) {
      double p = rec[c];
      double gg = li == c ? p - (double)1.0f : p;
      double hh = factor * p * ((double)1.0f - p);
      g[(size_t)c * n + i] = (float)gg;
      h[(size_t)c * n + i] = (float)hh;
    }
  }
}
// objective/multiclass.rs:156-158
static double multiclass_boost_from_score(const float* label, int64_t n, int K, int k) {
  std::vector<double> probs(K, 0.0);
  for (int64_t i = 0; i < n; ++i) probs[(int)label[i]] += 1.0;
  double p = probs[k] / (double)n;
  return std::log(std::max(K_EPSILON, p));
}

// ---------------------------------------------------------------- lambdarank
// dcg.rs:27-121: default label gains 2^i - 1 (i < 31), discounts 1/log2(2+i)
struct Dcg {
  std::vector<double> label_gain, discount;
  Dcg() {
    label_gain.push_back(0.0);
    for (int i = 1; i < 31; ++i) label_gain.push_back((double)((1 << i) - 1));
    for (int i = 0; i < 10000; ++i) discount.push_back(1.0 / std::log2(2.0 + (double)i));
  }
  // dcg.rs:68-84 CalMaxDCGAtK
  double max_dcg_at_k(int k, const float* label, int cnt) const {
    std::vector<int> c(label_gain.size(), 0);
    for (int i = 0; i < cnt; ++i) c[(int)label[i]]++;
    int top = (int)label_gain.size() - 1;
    double ret = 0.0;
    for (int j = 0; j < std::min(k, cnt); ++j) {
      while (top > 0 && c[top] <= 0) top--;
      if (top < 0) break;
      ret += discount[j] * label_gain[top];
      c[top]--;
    }
    return ret;
  }
  // dcg.rs:105-118 CalDCG at one k
  double dcg_at_k(int k, const float* label, const double* score, int cnt) const {
    std::vector<int> idx(cnt);
    for (int i = 0; i < cnt; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return score[a] > score[b]; });
    double cur = 0.0;
    for (int j = 0; j < std::min(k, cnt); ++j) cur += label_gain[(int)label[idx[j]]] * discount[j];
    return cur;
  }
};

// objective/rank.rs:129-257 LambdarankNdcg (no position bias, no weights)
struct Lambdarank {
  double sigmoid;
  bool norm;
  int trunc;
  Dcg dcg;
  std::vector<double> inv_max_dcg, table;
  double min_in = -50.0, max_in = 50.0, idx_factor = 0.0;
  static const int kBins = 1024 * 1024;
  Lambdarank(double sig, bool nrm, int tr, const float* label, const std::vector<int>& qb) : sigmoid(sig), norm(nrm), trunc(tr) {
    int k = std::max(tr, 0);
    for (size_t q = 0; q + 1 < qb.size(); ++q) {
      double v = dcg.max_dcg_at_k(k, label + qb[q], qb[q + 1] - qb[q]);
      inv_max_dcg.push_back(v > 0.0 ? 1.0 / v : v);
    }
    // rank.rs:166-179 ConstructSigmoidTable
    min_in = min_in / sigmoid / 2.0;
    max_in = -min_in;
    idx_factor = (double)kBins / (max_in - min_in);
    table.resize(kBins);
    for (int i = 0; i < kBins; ++i) {
      double sc = (double)i / idx_factor + min_in;
      table[i] = 1.0 / (1.0 + std::exp(sc * sigmoid));
    }
  }
  double get_sigmoid(double s) const {
    if (s <= min_in) return table[0];
    if (s >= max_in) return table[kBins - 1];
    return table[(size_t)((s - min_in) * idx_factor)];
  }
  // rank.rs:192-257 one_query
  void one_query(int q, const float* label, const double* score, int cnt, float* lambdas, float* hess) const {
    double imd = inv_max_dcg[q];
    for (int i = 0; i < cnt; ++i) { lambdas[i] = 0.0f; hess[i] = 0.0f; }
    if (cnt == 0) return;
    std::vector<int> sorted(cnt);
    for (int i = 0; i < cnt; ++i) sorted[i] = i;
    std::stable_sort(sorted.begin(), sorted.end(), [&](int a, int b) { return score[a] > score[b]; });
    double best = score[sorted[0]];
    int worst_idx = cnt - 1;
    if (worst_idx > 0 && score[sorted[worst_idx]] == K_MIN_SCORE) worst_idx--;
    double worst = score[sorted[worst_idx]];
    double sum_lambdas = 0.0;
    for (int i = 0; i + 1 < cnt && (long)i < (long)trunc; ++i) {
      if (score[sorted[i]] == K_MIN_SCORE) continue;
      for (int j = i + 1; j < cnt; ++j) {
        if (score[sorted[j]] == K_MIN_SCORE) continue;
        if (label[sorted[i]] == label[sorted[j]]) continue;
        int hr = label[sorted[i]] > label[sorted[j]] ? i : j;
        int lr = label[sorted[i]] > label[sorted[j]] ? j : i;
        int high = sorted[hr], low = sorted[lr];
        double hs = score[high], ls = score[low];
        double hg = dcg.label_gain[(int)label[high]], lg = dcg.label_gain[(int)label[low]];
        double hd = dcg.discount[hr], ld = dcg.discount[lr];
        double ds = hs - ls;
        double gap = hg - lg;
        double pd = std::fabs(hd - ld);
        double dp = gap * pd * imd;
        if (norm && best != worst) dp /= (double)0.01f + std::fabs(ds);
        double pl = get_sigmoid(ds);
        double ph = pl * (1.0 - pl);
        pl *= -sigmoid * dp;
        ph *= sigmoid * sigmoid * dp;
        lambdas[low] -= (float)pl;
        hess[low] += (float)ph;
        lambdas[high] += (float)pl;
        hess[high] += (float)ph;
        sum_lambdas -= 2.0 * pl;
      }
    }
    if (norm && sum_lambdas > 0.0) {
      double f = std::log2(1.0 + sum_lambdas) / sum_lambdas;
      for (int i = 0; i < cnt; ++i) {
        lambdas[i] = (float)((double)lambdas[i] * f);
        hess[i] = (float)((double)hess[i] * f);
      }
    }
  }
  void gradients(const double* score, const float* label, const std::vector<int>& qb, float* g, float* h) const {
    for (size_t q = 0; q + 1 < qb.size(); ++q)
      one_query((int)q, label + qb[q], score + qb[q], qb[q + 1] - qb[q], g + qb[q], h + qb[q]);
  }
};

} // namespace orc

using namespace orc;

// ---------------------------------------------------------------- booster
// sample_strategy.rs:186-214, 278-318: row bagging (no balanced fractions,
// no query bagging, no GOSS). One LCG per 1024-row block seeded
// bagging_seed + block; a row is in the bag when its draw is below the
// fraction.
// threading.rs block_info_force_size(num_threads, n, min_size): at most
// num_threads blocks of a multiple of min_size rows (one block of n rows when
// a single block results).
static void block_info_force_size(int num_threads, int64_t n, int64_t min_size, int64_t* nblock, int64_t* block) {
  *nblock = std::min<int64_t>(std::max(num_threads, 1), (n + min_size - 1) / min_size);
  *block = n;
  if (*nblock > 1) {
    *block = (n + *nblock - 1) / *nblock;
    *block = (*block + min_size - 1) / min_size * min_size;
  }
  if (*nblock < 1) *nblock = 1;
}

struct Bagger {
  bool on = false, subset = false, need_re_bagging = false, balanced = false;
  // GOSS (sample_strategy.rs:319-464): blocks of block_info_force_size(num_threads, n, 1024)
  bool goss = false;
  double top_rate = 0, other_rate = 0, learning_rate = 0.1;
  int num_threads = 1, K = 1;
  double fraction = 1.0, pos_fraction = 1.0, neg_fraction = 1.0;
  std::vector<float> label;
  int freq = 0;
  int64_t n = 0, bag_cnt = 0;
  std::vector<Random> rands;
  std::vector<uint32_t> in_bag, out_of_bag;
  // num_pos: the objective's num_positive_data (binary: labels > 0; 0 otherwise)
  void setup(const Params& p, int64_t rows, int num_feature_groups, int64_t num_pos, const float* lab) {
    n = rows;
    bag_cnt = rows;
    if (p.goss) {
      goss = on = true;
      top_rate = p.top_rate;
      other_rate = p.other_rate;
      learning_rate = p.learning_rate;
      num_threads = std::max(1, p.num_threads);
      K = p.objective == 1 ? p.num_class : 1;
      for (int64_t b = 0; b < (n + 1023) / 1024; ++b) rands.emplace_back((int)((uint32_t)p.bagging_seed + (uint32_t)b));
      subset = top_rate + other_rate <= 0.5;
      return;
    }
    balanced = (p.pos_bagging_fraction < 1.0 || p.neg_bagging_fraction < 1.0) && num_pos > 0;
    if (!((p.bagging_fraction < 1.0 || balanced) && p.bagging_freq > 0)) return;
    on = true;
    fraction = p.bagging_fraction;
    pos_fraction = p.pos_bagging_fraction;
    neg_fraction = p.neg_bagging_fraction;
    freq = p.bagging_freq;
    if (balanced) {
      bag_cnt = (int64_t)((double)num_pos * pos_fraction) + (int64_t)((double)(n - num_pos) * neg_fraction);
      label.assign(lab, lab + n);
    } else {
      bag_cnt = (int64_t)(fraction * (double)n);
    }
    for (int64_t b = 0; b < (n + 1023) / 1024; ++b) rands.emplace_back((int)((uint32_t)p.bagging_seed + (uint32_t)b));
    const double rate = ((double)bag_cnt / (double)n) / (double)freq;
    subset = rate <= 0.5 && num_feature_groups < 100;
    need_re_bagging = true;
  }
  // GOSSStrategy::Helper over rows start .. start + cnt (sample_strategy.rs:395-442)
  void goss_block(int64_t start, int64_t cnt, float* grad, float* hess) {
    auto abs_gh = [&](int64_t row) {
      float s = 0.0f;
      for (int k = 0; k < K; ++k) s += std::fabs(grad[(size_t)k * n + row] * hess[(size_t)k * n + row]);
      return s;
    };
    std::vector<float> tmp(cnt);
    for (int64_t i = 0; i < cnt; ++i) tmp[i] = abs_gh(start + i);
    const int top_k = std::max((int)((double)cnt * top_rate), 1);
    const int other_k = (int)((double)cnt * other_rate);
    // array_args.rs:65-79 ArgMaxAtK: position top_k - 1 of the descending order
    std::nth_element(tmp.begin(), tmp.begin() + (top_k - 1), tmp.end(), std::greater<float>());
    const float threshold = tmp[top_k - 1];
    const float multiply = (float)(int)(cnt - top_k) / (float)other_k;
    int big = 0, left = 0;
    for (int64_t i = 0; i < cnt; ++i) {
      const int64_t cur = start + i;
      const float g = abs_gh(cur);
      if (g >= threshold) {
        in_bag.push_back((uint32_t)cur);
        ++left;
        ++big;
      } else {
        const int sampled = left - big;
        const int rest_need = other_k - sampled;
        const int rest_all = (int)(cnt - i) - (top_k - big);
        const double prob = (double)rest_need / (double)rest_all;
        if ((double)rands[cur / 1024].next_float() < prob) {
          in_bag.push_back((uint32_t)cur);
          ++left;
          for (int k = 0; k < K; ++k) {
            grad[(size_t)k * n + cur] *= multiply;
            hess[(size_t)k * n + cur] *= multiply;
          }
        } else {
          out_of_bag.push_back((uint32_t)cur);
        }
      }
    }
  }
  bool bagging_goss(int iter, float* grad, float* hess) {
    bag_cnt = n;
    if ((int64_t)iter < (int64_t)((double)1.0f / learning_rate)) return false;
    int64_t nblock, block;
    block_info_force_size(num_threads, n, 1024, &nblock, &block);
    in_bag.clear();
    out_of_bag.clear();
    for (int64_t b = 0; b < nblock; ++b) {
      const int64_t start = b * block;
      if (start >= n) break;
      goss_block(start, std::min(block, n - start), grad, hess);
    }
    bag_cnt = (int64_t)in_bag.size();
    return true;
  }
  // bagging_by_query: the query boundaries (sample_strategy.rs:239-276)
  bool by_query = false;
  std::vector<int> qb;
  bool bagging(int iter) {
    if (!on || !((bag_cnt < n && iter % freq == 0) || need_re_bagging)) return false;
    need_re_bagging = false;
    in_bag.clear();
    out_of_bag.clear();
    if (by_query) {
      // query q is drawn by the generator of its 1024-query block; the bag is
      // the rows of the drawn queries in order
      const int nq = (int)qb.size() - 1;
      std::vector<char> in((size_t)n, 0);
      for (int q = 0; q < nq; ++q)
        if ((double)rands[q / 1024].next_float() < fraction)
          for (int r = qb[q]; r < qb[q + 1]; ++r) in[r] = 1;
      for (int64_t i = 0; i < n; ++i) (in[i] ? in_bag : out_of_bag).push_back((uint32_t)i);
      bag_cnt = (int64_t)in_bag.size();
      return true;
    }
    for (int64_t i = 0; i < n; ++i) {
      const double r = (double)rands[i / 1024].next_float();
      const double frac = balanced ? (label[i] > 0.0f ? pos_fraction : neg_fraction) : fraction;
      if (r < frac) in_bag.push_back((uint32_t)i);
      else out_of_bag.push_back((uint32_t)i);
    }
    bag_cnt = (int64_t)in_bag.size();
    return true;
  }
};

struct OrcBooster {
  Dataset* ds;
  Params prm;
  Learner* learner;
  int K;
  std::vector<double> scores;
  std::vector<float> grad, hess;
  std::vector<Tree> models;
  std::vector<double> init_scores;
  int iter = 0;
  std::vector<std::vector<TraceEntry>> traces;
  Lambdarank* rank = nullptr;
  Bagger bagger;
};

// boosting.rs:957-1139 (gbdt, no bagging, built-in objective)
static int booster_train_one_iter(OrcBooster* b) {
  int64_t n = b->ds->n;
  int K = b->K;
  std::vector<double> init(K, 0.0);
  for (int k = 0; k < K; ++k) {
    // boosting.rs:484-511
    if (b->models.empty() && b->prm.boost_from_average) {
      double v = 0.0;
      if (b->prm.objective == 2) v = 0.0;
      else if (b->prm.objective == 0) v = binary_boost_from_score(b->ds->label.data(), n, b->prm.sigmoid);
      else if (b->prm.objective == 1) v = multiclass_boost_from_score(b->ds->label.data(), n, K, k);
      if (std::fabs(v) > K_EPSILON) {
        for (int64_t i = 0; i < n; ++i) b->scores[(size_t)k * n + i] += v;
        init[k] = v;
      }
    }
  }
  if (!b->bagger.goss && b->bagger.bagging(b->iter)) {
    b->learner->bag = b->bagger.in_bag;
    b->learner->use_bag = true;
    b->learner->bag_subset = b->bagger.subset;
  }
  if (b->prm.objective == 0) binary_gradients(b->scores.data(), b->ds->label.data(), n, b->prm.sigmoid, b->grad.data(), b->hess.data());
  else if (b->prm.objective == 1) multiclass_gradients(b->scores.data(), b->ds->label.data(), n, K, b->grad.data(), b->hess.data());
  else b->rank->gradients(b->scores.data(), b->ds->label.data(), b->ds->query_boundaries, b->grad.data(), b->hess.data());
  // boosting.rs:1029-1035: the sampler after the gradients (GOSS rescales them)
  if (b->bagger.goss && b->bagger.bagging_goss(b->iter, b->grad.data(), b->hess.data())) {
    b->learner->bag = b->bagger.in_bag;
    b->learner->use_bag = true;
    b->learner->bag_subset = b->bagger.subset;
  }
  bool cont = false;
  for (int k = 0; k < K; ++k) {
    Tree t = b->learner->train(b->grad.data() + (size_t)k * n, b->hess.data() + (size_t)k * n);
    b->traces.push_back(b->learner->trace);
    if (t.num_leaves > 1) {
      cont = true;
      t.shrink(b->prm.learning_rate);
      b->learner->add_leaf_outputs(t, b->scores.data() + (size_t)k * n);
      // boosting.rs:1395-1398: out-of-bag rows by the binned tree
      if (b->bagger.on && b->bagger.bag_cnt < n)
        for (uint32_t r : b->bagger.out_of_bag) b->scores[(size_t)k * n + r] += t.leaf_value[b->learner->leaf_binned(t, r)];
      if (std::fabs(init[k]) > K_EPSILON) t.add_bias(init[k]);
    } else if (b->models.size() < (size_t)K) {
      t.as_constant(init[k], (int)n);
    } else {
      t.as_constant(0.0, (int)n);
    }
    b->models.push_back(t);
  }
  if (!cont) {
    if (b->models.size() > (size_t)K) for (int k = 0; k < K; ++k) b->models.pop_back();
    return 1;
  }
  b->iter++;
  return 0;
}

extern "C" {

int orc_version() { return 1; }
// Host threads of the CPU engine's loops (results are the same at any count).
void orc_set_num_threads(int n) { g_threads = std::max(1, n); }
void orc_block_info_force_size(int num_threads, int64_t n, int64_t min_size, int64_t* nblock, int64_t* block) {
  block_info_force_size(num_threads, n, min_size, nblock, block);
}

void* orc_dataset_create(const float* X, int64_t n, int ncol, const float* label, const char* params) {
  Params p;
  if (!parse_params(params, p)) return nullptr;
  Dataset* ds = build_dataset(X, n, ncol, p);
  if (label) ds->label.assign(label, label + n);
  return ds;
}
void orc_dataset_free(void* d) { delete (Dataset*)d; }
int orc_dataset_num_features(void* d) { return (int)((Dataset*)d)->used.size(); }
int orc_dataset_num_groups(void* d) { return ((Dataset*)d)->num_groups; }
// Bin mapper of inner feature f: returns num_bin; fills the scalar fields and,
// when `upper` is given, num_bin upper bounds (categories for categorical).
int orc_dataset_feature(void* d, int f, int* real, int* missing_type, int* default_bin, int* mfb, int* bin_type,
                        double* upper) {
  Dataset* ds = (Dataset*)d;
  const BinMapper& m = ds->mappers[ds->used[f]];
  *real = ds->used[f];
  *missing_type = m.missing_type;
  *default_bin = (int)m.default_bin;
  *mfb = (int)m.most_freq_bin;
  *bin_type = m.bin_type;
  if (upper)
    for (int i = 0; i < m.num_bin; ++i) upper[i] = m.bin_type == BT_CAT ? (double)m.bin_2_categorical[i] : m.bin_upper_bound[i];
  return m.num_bin;
}
// Bins of every row, n x F row-major (u16 to cover wide features).
void orc_dataset_bins(void* d, uint16_t* out) {
  Dataset* ds = (Dataset*)d;
  size_t F = ds->used.size();
  for (int64_t r = 0; r < ds->n; ++r)
    for (size_t f = 0; f < F; ++f) out[(size_t)r * F + f] = (uint16_t)ds->bin(r, (int)f);
}

void* orc_booster_create(void* d, const char* params) {
  Params p;
  if (!parse_params(params, p)) return nullptr;
  OrcBooster* b = new OrcBooster();
  b->ds = (Dataset*)d;
  b->prm = p;
  b->K = p.objective == 1 ? p.num_class : 1;
  b->learner = new Learner(b->ds, p);
  if (p.objective == 2) {
    if (b->ds->query_boundaries.empty()) { delete b->learner; delete b; return nullptr; }
    b->rank = new Lambdarank(p.sigmoid, p.lambdarank_norm, p.lambdarank_truncation_level, b->ds->label.data(),
                             b->ds->query_boundaries);
  }
  b->scores.assign((size_t)b->K * b->ds->n, 0.0);
  b->grad.assign((size_t)b->K * b->ds->n, 0.0f);
  b->hess.assign((size_t)b->K * b->ds->n, 0.0f);
  int64_t num_pos = 0;
  if (p.objective == 0)
    for (float l : b->ds->label) num_pos += l > 0.0f ? 1 : 0;
  b->bagger.setup(p, b->ds->n, b->ds->num_groups, num_pos, b->ds->label.data());
  if (p.bagging_by_query && b->bagger.on && !b->bagger.goss) {
    // sample_strategy.rs:239-276: whole queries are drawn (no query data: no queries)
    b->bagger.by_query = true;
    b->bagger.qb = b->ds->query_boundaries.empty() ? std::vector<int>{0} : b->ds->query_boundaries;
  }
  return b;
}
// The current bag (in-bag rows ascending); returns its size, or -1 without bagging.
int64_t orc_booster_bag(void* p, uint32_t* out) {
  OrcBooster* b = (OrcBooster*)p;
  if (!b->bagger.on) return -1;
  if (out) std::memcpy(out, b->bagger.in_bag.data(), b->bagger.in_bag.size() * 4);
  return (int64_t)b->bagger.in_bag.size();
}
void orc_booster_free(void* p) {
  OrcBooster* b = (OrcBooster*)p;
  delete b->learner;
  delete b->rank;
  delete b;
}

// A prediction-only booster from model text (io.rs:156-343 header and tree
// blocks, tree.rs:869-1047 fields): num_tree_per_iteration (or num_class) and
// each tree's num_leaves, num_cat, split_feature, threshold, decision_type,
// children, leaf_value and categorical bitsets. Numbers are read with strtod
// (correctly rounded, as parse_f64). Returns null on a malformed text.
void* orc_model_load(const char* text) {
  std::vector<std::string> lines;
  std::string cur;
  for (const char* q = text; *q; ++q) {
    if (*q == '\n') { lines.push_back(cur); cur.clear(); }
    else if (*q != '\r') cur += *q;
  }
  lines.push_back(cur);
  auto vals = [](const std::string& v) {
    std::vector<double> out;
    const char* q = v.c_str();
    while (*q) {
      while (*q == ' ') ++q;
      if (!*q) break;
      char* e;
      out.push_back(std::strtod(q, &e));
      if (e == q) return std::vector<double>();
      q = e;
    }
    return out;
  };
  int K = -1, num_class = 1;
  size_t i = 0;
  for (; i < lines.size() && lines[i].rfind("Tree=", 0) != 0; ++i) {
    if (lines[i].rfind("num_tree_per_iteration=", 0) == 0) K = std::atoi(lines[i].c_str() + 23);
    if (lines[i].rfind("num_class=", 0) == 0) num_class = std::atoi(lines[i].c_str() + 10);
  }
  if (K < 1) K = num_class;
  OrcBooster* b = new OrcBooster();
  b->ds = nullptr;
  b->learner = nullptr;
  b->K = K;
  while (i < lines.size() && lines[i].rfind("Tree=", 0) == 0) {
    std::map<std::string, std::string> kv;
    for (++i; i < lines.size() && !lines[i].empty(); ++i) {
      const size_t eq = lines[i].find('=');
      if (eq != std::string::npos) kv[lines[i].substr(0, eq)] = lines[i].substr(eq + 1);
    }
    while (i < lines.size() && lines[i].empty()) ++i;
    Tree t;
    const int n = std::atoi(kv["num_leaves"].c_str());
    if (n < 1) { delete b; return nullptr; }
    t.num_leaves = n;
    t.num_cat = std::atoi(kv["num_cat"].c_str());
    const auto lv = vals(kv["leaf_value"]);
    if ((int)lv.size() != n) { delete b; return nullptr; }
    t.leaf_value = lv;
    if (n > 1) {
      const auto sf = vals(kv["split_feature"]), th = vals(kv["threshold"]), dt = vals(kv["decision_type"]),
                 lc = vals(kv["left_child"]), rc = vals(kv["right_child"]);
      if ((int)sf.size() != n - 1 || (int)th.size() != n - 1 || (int)dt.size() != n - 1 || (int)lc.size() != n - 1 ||
          (int)rc.size() != n - 1) { delete b; return nullptr; }
      for (int k = 0; k < n - 1; ++k) {
        t.split_feature.push_back((int)sf[k]);
        t.threshold.push_back(th[k]);
        t.decision_type.push_back((int8_t)(int)dt[k]);
        t.left_child.push_back((int)lc[k]);
        t.right_child.push_back((int)rc[k]);
      }
      if (t.num_cat > 0) {
        const auto cb = vals(kv["cat_boundaries"]), ct = vals(kv["cat_threshold"]);
        t.cat_boundaries.assign(cb.begin(), cb.end());
        for (double x : ct) t.cat_threshold.push_back((uint32_t)(long long)x);
      }
    }
    b->models.push_back(std::move(t));
  }
  if (b->models.empty()) { delete b; return nullptr; }
  return b;
}
int orc_booster_train_one_iter(void* p) { return booster_train_one_iter((OrcBooster*)p); }
// One GOSS draw (sample_strategy.rs:320-442) on caller gradients g, h [K][n]
// (rescaled in place): a fresh sampler of the given seed, thread count and
// rates at an iteration past the warm-up. Writes the in-bag rows in order and
// returns their count.
int64_t orc_goss(float* g, float* h, int64_t n, int K, int num_threads, double top_rate, double other_rate,
                 int seed, uint32_t* in_bag) {
  Params p;
  p.goss = 1;
  p.top_rate = top_rate;
  p.other_rate = other_rate;
  p.num_threads = num_threads;
  p.bagging_seed = seed;
  p.learning_rate = 1.0;
  p.objective = K > 1 ? 1 : 0;
  p.num_class = K;
  Bagger bg;
  bg.setup(p, n, 1, 0, nullptr);
  bg.bagging_goss(1, g, h);
  std::memcpy(in_bag, bg.in_bag.data(), bg.in_bag.size() * 4);
  return (int64_t)bg.in_bag.size();
}
int orc_booster_num_trees(void* p) { return (int)((OrcBooster*)p)->models.size(); }
void orc_booster_scores(void* p, double* out) {
  OrcBooster* b = (OrcBooster*)p;
  std::memcpy(out, b->scores.data(), b->scores.size() * 8);
}
void orc_booster_gradients(void* p, float* g, float* h) {
  OrcBooster* b = (OrcBooster*)p;
  std::memcpy(g, b->grad.data(), b->grad.size() * 4);
  std::memcpy(h, b->hess.data(), b->hess.size() * 4);
}
int orc_tree_num_leaves(void* p, int t) { return ((OrcBooster*)p)->models[t].num_leaves; }
// Tree arrays: node arrays have num_leaves-1 entries, leaf arrays num_leaves.
void orc_tree_arrays(void* p, int t, int* split_feature, uint32_t* threshold_in_bin, double* threshold,
                     int8_t* decision_type, int* left_child, int* right_child, float* split_gain, double* leaf_value,
                     int* leaf_count, double* leaf_weight, double* internal_value) {
  const Tree& tr = ((OrcBooster*)p)->models[t];
  int ni = tr.num_leaves - 1;
  for (int i = 0; i < ni; ++i) {
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
}
// The split trace of tree t: (leaf, real feature, threshold bin, gain).
int orc_tree_trace(void* p, int t, int* leaf, int* feature, uint32_t* thr, double* gain) {
  const auto& tr = ((OrcBooster*)p)->traces[t];
  for (size_t i = 0; i < tr.size(); ++i) { leaf[i] = tr[i].leaf; feature[i] = tr[i].feature; thr[i] = tr[i].threshold; gain[i] = tr[i].gain; }
  return (int)tr.size();
}

// Train one tree from caller gradients on the booster's dataset (no score
// update): the stage-level learner check.
int orc_learner_train(void* p, const float* g, const float* h) {
  OrcBooster* b = (OrcBooster*)p;
  Tree t = b->learner->train(g, h);
  b->models.push_back(t);
  b->traces.push_back(b->learner->trace);
  return t.num_leaves;
}
// Leaf of each row after the last orc_learner_train (partition order).
void orc_learner_leaf_of_row(void* p, int* out) {
  OrcBooster* b = (OrcBooster*)p;
  Learner* l = b->learner;
  for (int leaf = 0; leaf < (int)l->leaf_count.size(); ++leaf)
    for (int64_t k = 0; k < l->leaf_count[leaf]; ++k) out[l->indices[l->leaf_begin[leaf] + k]] = leaf;
}
// The last tree's root sums (learner/mod.rs:611-628): gradient, hessian, packed integer sum.
void orc_learner_root_sums(void* p, double* sg, double* sh, int64_t* is) {
  const LeafSplits& r = ((OrcBooster*)p)->learner->root_sums;
  *sg = r.sum_gradients;
  *sh = r.sum_hessians;
  *is = r.int_sum;
}
void orc_learner_quant_scales(void* p, double* gs, double* hs) {
  OrcBooster* b = (OrcBooster*)p;
  *gs = b->learner->quant ? b->learner->quant->grad_scale : 0.0;
  *hs = b->learner->quant ? b->learner->quant->hess_scale : 0.0;
}
void orc_learner_quant_packed(void* p, int8_t* gq, int8_t* hq) {
  OrcBooster* b = (OrcBooster*)p;
  Discretizer* q = b->learner->quant;
  for (size_t i = 0; i < q->gq.size(); ++i) { gq[i] = q->gq[i]; hq[i] = q->hq[i]; }
}

// Raw prediction (predict.rs:231-339 / tree.rs:417-443): sum of leaf values
// over trees in model order, class-major output [rows][K].
void orc_booster_predict_raw(void* p, const float* X, int64_t rows, int ncol, double* out) {
  OrcBooster* b = (OrcBooster*)p;
  int K = b->K;
  std::vector<double> row(ncol);
  for (int64_t r = 0; r < rows; ++r) {
    for (int c = 0; c < ncol; ++c) {
      double v = (double)X[(size_t)r * ncol + c];
      if (!std::isnan(v) && std::fabs(v) <= K_ZERO_THRESHOLD) v = 0.0;
      row[c] = v;
    }
    for (int k = 0; k < K; ++k) out[(size_t)r * K + k] = 0.0;
    for (size_t m = 0; m < b->models.size(); ++m) {
      const Tree& t = b->models[m];
      out[(size_t)r * K + (m % K)] += t.leaf_value[t.get_leaf(row.data(), ncol)];
    }
  }
}

// predict.rs:193-202 window and :231-339 with early stopping (:155-171):
// iterations start .. start + num in order, each class's tree added to its
// raw score; every freq iterations (freq > 0) a row whose margin exceeds the
// threshold stops.
void orc_booster_predict_window(void* p, const float* X, int64_t rows, int ncol, int start_iteration,
                                int num_iteration, int freq, double margin, double* out) {
  OrcBooster* b = (OrcBooster*)p;
  const int K = b->K;
  const int total = (int)b->models.size() / K;
  const int start = std::min(std::max(start_iteration, 0), total);
  const int num = num_iteration > 0 ? std::min(num_iteration, total - start) : total - start;
  std::vector<double> row(ncol);
  for (int64_t r = 0; r < rows; ++r) {
    for (int c = 0; c < ncol; ++c) {
      double v = (double)X[(size_t)r * ncol + c];
      if (!std::isnan(v) && std::fabs(v) <= K_ZERO_THRESHOLD) v = 0.0;
      row[c] = v;
    }
    double* raw = out + (size_t)r * K;
    for (int k = 0; k < K; ++k) raw[k] = 0.0;
    for (int it = 0; it < num; ++it) {
      for (int k = 0; k < K; ++k) {
        const Tree& t = b->models[(size_t)(start + it) * K + k];
        raw[k] += t.leaf_value[t.get_leaf(row.data(), ncol)];
      }
      if (freq > 0 && (it + 1) % freq == 0) {
        double m;
        if (K > 1) {
          double x = -INFINITY, y = -INFINITY;
          for (int k = 0; k < K; ++k) {
            if (raw[k] > x) { y = x; x = raw[k]; }
            else if (raw[k] > y) y = raw[k];
          }
          m = x - y;
        } else {
          m = 2.0 * std::fabs(raw[0]);
        }
        if (m > margin) break;
      }
    }
  }
}

// --------------------------- stage functions (same layouts as the HIP engine)
// Feature metas are passed as int arrays of length F:
//   num_bin, missing_type, default_bin, most_freq_bin, bin_type.

// FP64 histogram of rows idx[0..m) in the given order; bins are n x F
// row-major u8; output [F][B][2], most frequent bins left at zero.
void orc_hist_f64(const uint8_t* bins, int F, const int* mfb, int B, const uint32_t* idx, int64_t m, const float* g,
                  const float* h, double* out) {
  std::memset(out, 0, sizeof(double) * (size_t)F * B * 2);
  for (int64_t k = 0; k < m; ++k) {
    uint32_t r = idx[k];
    for (int f = 0; f < F; ++f) {
      uint32_t b = bins[(size_t)r * F + f];
      if ((int)b == mfb[f]) continue;
      out[((size_t)f * B + b) * 2] += (double)g[r];
      out[((size_t)f * B + b) * 2 + 1] += (double)h[r];
    }
  }
}
// Packed-integer histogram (wrapping i64 sums of (gq << 32) | u8(hq)).
void orc_hist_int(const uint8_t* bins, int F, const int* mfb, int B, const uint32_t* idx, int64_t m,
                  const int8_t* gq, const int8_t* hq, int64_t* out) {
  std::memset(out, 0, sizeof(int64_t) * (size_t)F * B);
  for (int64_t k = 0; k < m; ++k) {
    uint32_t r = idx[k];
    int64_t v = (int64_t)(((uint64_t)(int64_t)gq[r] << 32) | (uint64_t)(uint8_t)hq[r]);
    for (int f = 0; f < F; ++f) {
      uint32_t b = bins[(size_t)r * F + f];
      if ((int)b == mfb[f]) continue;
      out[(size_t)f * B + b] = wadd(out[(size_t)f * B + b], v);
    }
  }
}

static SplitParams stage_params(const double* dp, const int* ip) {
  // dp: lambda_l1, lambda_l2, max_delta_step, path_smooth, min_sum_hessian, min_gain_to_split, cat_l2, cat_smooth
  // ip: min_data_in_leaf, max_cat_to_onehot, max_cat_threshold, min_data_per_group
  return SplitParams{dp[0], dp[1], dp[2], dp[3], ip[0], dp[4], dp[5], ip[1], ip[2], dp[6], dp[7], ip[3]};
}

// histogram.rs:42-85: the most-frequent-bin fix on a bin-indexed histogram
// ([num_bin][2] FP64, or [num_bin] packed int64), on its own for tests.
void orc_hist_fix(int num_bin, int mfb, double sg, double sh, double* hist) {
  FeatureMeta m{num_bin, 0, mfb == 0 ? 1 : 0, 0u, (uint32_t)mfb, 0, 1.0};
  hist_fix(m, sg, sh, hist);
}
void orc_hist_fix_int(int num_bin, int mfb, int64_t s, int64_t* hist) {
  FeatureMeta m{num_bin, 0, mfb == 0 ? 1 : 0, 0u, (uint32_t)mfb, 0, 1.0};
  hist_fix_int(m, s, hist);
}

// Split record: out_d = {gain, left_output, right_output, lsg, lsh, rsg, rsh},
// out_i = {threshold, default_left, left_count, right_count, splittable}.
// `hist` is the bin-indexed [B][2] histogram of one feature; the most
// frequent bin is fixed from (sg, sh) first, as the learner does.
void orc_find_best_threshold(double* hist, int num_bin, int missing_type, int default_bin, int mfb, int bin_type,
                             const double* dp, const int* ip, double sg, double sh, int nd, double parent_output,
                             double* out_d, int* out_i, int* out_cat) {
  FeatureMeta m{num_bin, missing_type, mfb == 0 ? 1 : 0, (uint32_t)default_bin, (uint32_t)mfb, bin_type, 1.0};
  SplitParams sp = stage_params(dp, ip);
  hist_fix(m, sg, sh, hist);
  SplitInfo s;
  bool ok = find_best_threshold(hist + 2 * m.offset, m, sp, sg, sh, nd, parent_output, s);
  out_d[0] = s.gain; out_d[1] = s.left_output; out_d[2] = s.right_output;
  out_d[3] = s.left_sum_gradient; out_d[4] = s.left_sum_hessian; out_d[5] = s.right_sum_gradient; out_d[6] = s.right_sum_hessian;
  out_i[0] = (int)s.threshold; out_i[1] = s.default_left; out_i[2] = s.left_count; out_i[3] = s.right_count; out_i[4] = ok;
  // out_cat: [n, bins sent left in the search's order] (num_bin + 1 ints)
  out_cat[0] = ok && bin_type == BT_CAT ? (int)s.cat_threshold.size() : 0;
  for (int i = 0; i < out_cat[0]; ++i) out_cat[1 + i] = (int)s.cat_threshold[i];
}
void orc_find_best_threshold_int(int64_t* hist, int num_bin, int missing_type, int default_bin, int mfb, int bin_type,
                                 const double* dp, const int* ip, int64_t int_sum, double gscale, double hscale,
                                 int nd, double parent_output, double* out_d, int* out_i, int64_t* out_gh,
                                 int* out_cat) {
  FeatureMeta m{num_bin, missing_type, mfb == 0 ? 1 : 0, (uint32_t)default_bin, (uint32_t)mfb, bin_type, 1.0};
  SplitParams sp = stage_params(dp, ip);
  hist_fix_int(m, int_sum, hist);
  SplitInfo s;
  bool ok = find_best_threshold_int(hist + m.offset, m, sp, IntSums{int_sum, gscale, hscale}, nd, parent_output, s);
  out_d[0] = s.gain; out_d[1] = s.left_output; out_d[2] = s.right_output;
  out_d[3] = s.left_sum_gradient; out_d[4] = s.left_sum_hessian; out_d[5] = s.right_sum_gradient; out_d[6] = s.right_sum_hessian;
  out_i[0] = (int)s.threshold; out_i[1] = s.default_left; out_i[2] = s.left_count; out_i[3] = s.right_count; out_i[4] = ok;
  out_gh[0] = s.left_sum_gh; out_gh[1] = s.right_sum_gh;
  out_cat[0] = ok && bin_type == BT_CAT ? (int)s.cat_threshold.size() : 0;
  for (int i = 0; i < out_cat[0]; ++i) out_cat[1 + i] = (int)s.cat_threshold[i];
}

void orc_binary_gradients(const double* s, const float* label, int64_t n, double sigmoid, float* g, float* h) {
  binary_gradients(s, label, n, sigmoid, g, h);
}
void orc_multiclass_gradients(const double* s, const float* label, int64_t n, int K, float* g, float* h) {
  multiclass_gradients(s, label, n, K, g, h);
}

// quant.rs discretize with num_threads = 1 tables; returns the start offset.
void orc_discretize(const float* g, const float* h, int64_t n, int bins, int seed, int stochastic, int num_calls,
                    int8_t* gq, int8_t* hq, double* scales) {
  Discretizer d(bins, seed, stochastic != 0, n, 1);
  for (int c = 0; c < num_calls; ++c) d.discretize(g, h, n);
  for (int64_t i = 0; i < n; ++i) { gq[i] = d.gq[i]; hq[i] = d.hq[i]; }
  scales[0] = d.grad_scale;
  scales[1] = d.hess_scale;
}
// The per-row random tables of a one-thread discretizer and the first
// `count` start offsets its generator draws (quant.rs:45-79, 114).
void orc_quant_tables(int64_t n, int seed, double* grv, double* hrv, int count, int* starts) {
  Discretizer d(4, seed, true, n, 1);
  for (int64_t i = 0; i < n; ++i) { grv[i] = d.grv[i]; hrv[i] = d.hrv[i]; }
  for (int c = 0; c < count; ++c) starts[c] = d.start_eng.uniform_int(0, d.start_max);
}
int orc_to_i8(double x) { return (int)to_i8(x); }
uint32_t orc_mt19937_nth(uint32_t seed, int nth) {
  Mt19937 m(seed);
  uint32_t v = 0;
  for (int i = 0; i < nth; ++i) v = m.next_u32();
  return v;
}
int orc_random_sample(int seed, int n, int k, int* out) {
  Random r(seed);
  auto s = r.sample(n, k);
  for (size_t i = 0; i < s.size(); ++i) out[i] = s[i];
  return (int)s.size();
}
// Bin mapper from raw sampled values (binning.rs:387-472) for known-answer tests.
int orc_find_bin(const double* vals, int nv, int64_t total, int bin_type, int max_bin, int min_data_in_bin,
                 int min_split_data, int pre_filter, double* upper, int* missing_type, int* default_bin, int* mfb,
                 int* is_trivial, int* b2c) {
  std::vector<double> v(vals, vals + nv);
  BinParams bp{max_bin, min_data_in_bin, min_split_data, pre_filter != 0, true, false};
  BinMapper m = find_bin(v, (size_t)total, bin_type, bp);
  for (int i = 0; i < m.num_bin; ++i) {
    if (m.bin_type == BT_NUM) upper[i] = m.bin_upper_bound[i];
    else b2c[i] = m.bin_2_categorical[i];
  }
  *missing_type = m.missing_type;
  *default_bin = (int)m.default_bin;
  *mfb = (int)m.most_freq_bin;
  *is_trivial = m.is_trivial;
  return m.num_bin;
}
int orc_value_to_bin_num(const double* upper, int num_bin, int missing_type, double v) {
  BinMapper m;
  m.num_bin = num_bin;
  m.missing_type = missing_type;
  m.bin_upper_bound.assign(upper, upper + num_bin);
  m.bin_type = BT_NUM;
  return (int)m.value_to_bin(v);
}

void orc_dataset_set_query(void* d, const int* qb, int nq) {
  ((Dataset*)d)->query_boundaries.assign(qb, qb + nq + 1);
}
// Lambdarank gradients of every query (rank.rs:63-108 without weights / position bias).
void orc_lambdarank_gradients(const double* score, const float* label, const int* qb, int nq, double sigmoid,
                              int norm, int trunc, float* g, float* h) {
  std::vector<int> b(qb, qb + nq + 1);
  Lambdarank lr(sigmoid, norm != 0, trunc, label, b);
  lr.gradients(score, label, b, g, h);
}
// metric.rs:551-600 NDCG@k averaged over queries (queries with max DCG 0 count as 1).
double orc_ndcg(const double* score, const float* label, const int* qb, int nq, int k) {
  Dcg d;
  double sum = 0.0;
  for (int q = 0; q < nq; ++q) {
    int cnt = qb[q + 1] - qb[q];
    double m = d.max_dcg_at_k(k, label + qb[q], cnt);
    if (m <= 0.0) { sum += 1.0; continue; }
    sum += d.dcg_at_k(k, label + qb[q], score + qb[q], cnt) / m;
  }
  return sum / (double)nq;
}
double orc_max_dcg_at_k(const float* label, int cnt, int k) { Dcg d; return d.max_dcg_at_k(k, label, cnt); }

// metric.rs:701-729 (stable order for ties; ties are grouped by threshold)
double orc_auc(const double* score, const float* label, int64_t n) {
  std::vector<int64_t> idx(n);
  for (int64_t i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return score[a] > score[b]; });
  double cur_pos = 0, sum_pos = 0, accum = 0, cur_neg = 0;
  double threshold = score[idx[0]];
  for (int64_t k = 0; k < n; ++k) {
    int64_t i = idx[k];
    double s = score[i];
    if (s != threshold) {
      threshold = s;
      accum += cur_neg * (cur_pos * 0.5 + sum_pos);
      sum_pos += cur_pos;
      cur_neg = 0;
      cur_pos = 0;
    }
    cur_neg += (double)(label[i] <= 0.0f);
    cur_pos += (double)(label[i] > 0.0f);
  }
  accum += cur_neg * (cur_pos * 0.5 + sum_pos);
  sum_pos += cur_pos;
  double sw = (double)n;
  return (sum_pos > 0.0 && sum_pos != sw) ? accum / (sum_pos * (sw - sum_pos)) : 1.0;
}

// metric.rs:602-625 binary logloss (unweighted)
double orc_binary_logloss(const double* raw, const float* label, int64_t n, double sigmoid) {
  double sum = 0.0;
  double eps_clamp = -(double)std::log(1e-15f);
  for (int64_t i = 0; i < n; ++i) {
    double p = 1.0 / (1.0 + std::exp(-sigmoid * raw[i]));
    double l;
    if (label[i] <= 0.0f) l = 1.0 - p > K_EPSILON ? -std::log(1.0 - p) : eps_clamp;
    else l = p > K_EPSILON ? -std::log(p) : eps_clamp;
    sum += l;
  }
  return sum / (double)n;
}

// ---------------------------------------------------------------- metrics
// Metric::new / Metric::auc (metric.rs:238-247, 701-729): the weight total in
// row order; documents by descending score (ties in index order, one of the
// orders the reference's unstable sort may produce); runs of equal scores.
double orc_metric_auc(const double* score, const float* label, const float* weight, int64_t n) {
  double sum_weights = (double)n;
  if (weight) {
    sum_weights = 0.0;
    for (int64_t i = 0; i < n; ++i) sum_weights += (double)weight[i];
  }
  std::vector<int64_t> idx(n);
  for (int64_t i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return score[a] > score[b]; });
  double cur_pos = 0.0, sum_pos = 0.0, accum = 0.0, cur_neg = 0.0;
  double threshold = score[idx[0]];
  for (int64_t k = 0; k < n; ++k) {
    const int64_t i = idx[k];
    const double s = score[i];
    if (s != threshold) {
      threshold = s;
      accum += cur_neg * (cur_pos * 0.5 + sum_pos);
      sum_pos += cur_pos;
      cur_neg = 0.0;
      cur_pos = 0.0;
    }
    const double w = weight ? (double)weight[i] : 1.0;
    cur_neg += (double)(label[i] <= 0.0f ? 1 : 0) * w;
    cur_pos += (double)(label[i] > 0.0f ? 1 : 0) * w;
  }
  accum += cur_neg * (cur_pos * 0.5 + sum_pos);
  sum_pos += cur_pos;
  return (sum_pos > 0.0 && sum_pos != sum_weights) ? accum / (sum_pos * (sum_weights - sum_pos)) : 1.0;
}

// metric.rs:602-625 (binary, binary.rs:164-166 sigmoid) and :629-653 with
// multiclass_loss :133-153 (softmax multiclass.rs:16-30). kind: 0
// binary_logloss, 1 binary_error, 2 multi_logloss, 3 multi_error. score is
// class-major [K][n].
double orc_metric_point(int kind, const double* score, const float* label, const float* weight, int64_t n, int K,
                        double sigmoid, int top_k) {
  double sum_weights = (double)n;
  if (weight) {
    sum_weights = 0.0;
    for (int64_t i = 0; i < n; ++i) sum_weights += (double)weight[i];
  }
  const double nle = -(double)std::log(1e-15f);
  double sum = 0.0;
  std::vector<double> raw(K), rec(K);
  for (int64_t i = 0; i < n; ++i) {
    double l;
    if (kind <= 1) {
      const double p = 1.0 / (1.0 + std::exp(-sigmoid * score[i]));
      if (kind == 0) {
        if (label[i] <= 0.0f) l = 1.0 - p > K_EPSILON ? -std::log(1.0 - p) : nle;
        else l = p > K_EPSILON ? -std::log(p) : nle;
      } else {
        l = p <= 0.5 ? (double)(label[i] > 0.0f ? 1 : 0) : (double)(label[i] <= 0.0f ? 1 : 0);
      }
    } else {
      for (int c = 0; c < K; ++c) raw[c] = score[(size_t)c * n + i];
      double wmax = raw[0];
      for (int c = 1; c < K; ++c) wmax = raw[c] < wmax ? wmax : raw[c];
      double wsum = 0.0;
      for (int c = 0; c < K; ++c) {
        rec[c] = std::exp(raw[c] - wmax);
        wsum += rec[c];
      }
      for (int c = 0; c < K; ++c) rec[c] /= wsum;
      const int k = (int)label[i];
      if (kind == 2) {
        l = rec[k] > K_EPSILON ? -std::log(rec[k]) : nle;
      } else {
        int larger = 0;
        l = 0.0;
        for (int c = 0; c < K; ++c) {
          if (rec[c] >= rec[k]) ++larger;
          if (larger > top_k) {
            l = 1.0;
            break;
          }
        }
      }
    }
    sum += weight ? l * (double)weight[i] : l;
  }
  return sum / sum_weights;
}

// Metric::rank_eval for NDCG (metric.rs:551-600) with DcgCalculator::dcg /
// max_dcg (dcg.rs:88-118): per query, DCG at every cut-off of the ranking by
// descending score (ties in index order), times 1 / max DCG (1 for a query
// without relevant documents); libgomp static chunks of queries over
// `threads`, chunk sums added in chunk order, divided by the query count.
void orc_metric_ndcg(const double* score, const float* label, const int* qb, int nq, const int* eval_at, int nk,
                     int threads, double* out) {
  Dcg d;
  threads = std::max(1, threads);
  std::vector<std::vector<double>> bufs(threads, std::vector<double>(nk, 0.0));
  std::vector<double> tmp(nk), inv(nk);
  const int qc = nq / threads, r = nq % threads;
  for (int t = 0; t < threads; ++t) {
    const int start = t * qc + std::min(t, r), len = qc + (t < r ? 1 : 0);
    for (int q = start; q < start + len; ++q) {
      const int s0 = qb[q], cnt = qb[q + 1] - s0;
      const float* lq = label + s0;
      // max DCG at every cut-off (dcg.rs:88-103)
      std::vector<int> c(d.label_gain.size(), 0);
      for (int i = 0; i < cnt; ++i) c[(int)lq[i]]++;
      int top = (int)d.label_gain.size() - 1, left = 0;
      double cur = 0.0;
      for (int m = 0; m < nk; ++m) {
        const int k = std::min(eval_at[m], cnt);
        for (int j = left; j < k; ++j) {
          while (top > 0 && c[top] <= 0) top--;
          cur += d.discount[j] * d.label_gain[top];
          c[top]--;
        }
        inv[m] = cur > 0.0 ? 1.0 / cur : -1.0;
        left = k;
      }
      if (inv[0] <= 0.0) {
        for (int m = 0; m < nk; ++m) bufs[t][m] += 1.0;
        continue;
      }
      // DCG at every cut-off (dcg.rs:105-118)
      std::vector<int> idx(cnt);
      for (int i = 0; i < cnt; ++i) idx[i] = i;
      std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return score[s0 + a] > score[s0 + b]; });
      cur = 0.0;
      left = 0;
      for (int m = 0; m < nk; ++m) {
        const int k = std::min(eval_at[m], cnt);
        for (int j = left; j < k; ++j) cur += d.label_gain[(int)lq[idx[j]]] * d.discount[j];
        tmp[m] = cur;
        left = k;
      }
      for (int m = 0; m < nk; ++m) bufs[t][m] += tmp[m] * inv[m];
    }
  }
  for (int m = 0; m < nk; ++m) {
    double v = 0.0;
    for (int t = 0; t < threads; ++t) v += bufs[t][m];
    out[m] = v / (double)nq;
  }
}

}  // extern "C"
