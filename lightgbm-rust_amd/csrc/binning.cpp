// Host-side binning: the BinMapper half of the boundary (it stays on the
// host; the GPU receives the packed bin matrix once).
//
// Restates crates/lgbm-core/src/binning.rs (find_bin_logged :398-517,
// greedy_find_bin :117-212, find_bin_with_zero_as_one_bin :215-276,
// value_to_bin :637-667), the row sampling of dataset.rs:393-397 with the
// upstream LCG (random.rs:45-70) and the dense column sampling of
// matrix.rs:198-214. Columns are binned in parallel on host threads; each
// column's result is independent of the thread count.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.h"

namespace lgbmhip {

namespace {

const double kZero = (double)1e-35f;  // consts.rs:15 kZeroThreshold
const double kSparse = 0.7;           // consts.rs:20

// random.rs:10-71
class Lcg {
 public:
  explicit Lcg(int seed) : x_((uint32_t)seed) {}
  int r16() { x_ = x_ * 214013u + 2531011u; return (int)((x_ >> 16) & 0x7FFF); }
  int r31() { x_ = x_ * 214013u + 2531011u; return (int)(x_ & 0x7FFFFFFF); }
  int next_int(int lo, int hi) { return r31() % (hi - lo) + lo; }
  float next_float() { return (float)r16() / 32768.0f; }
  std::vector<int> sample(int n, int k) {
    std::vector<int> out;
    if (k > n || k <= 0) return out;
    if (k == n) {
      out.resize(n);
      for (int i = 0; i < n; ++i) out[i] = i;
      return out;
    }
    if (k > 1 && (double)k > (double)n / std::log2((double)k)) {
      out.reserve(k);
      for (int i = 0; i < n; ++i) {
        double prob = (double)(k - (int)out.size()) / (double)(n - i);
        if ((double)next_float() < prob) out.push_back(i);
      }
      return out;
    }
    std::set<int> chosen;
    for (int r = n - k; r < n; ++r) {
      int v = next_int(0, r + 1);
      if (!chosen.insert(v).second) chosen.insert(r);
    }
    return std::vector<int>(chosen.begin(), chosen.end());
  }

 private:
  uint32_t x_;
};

inline double up(double a) { return std::nextafter(a, INFINITY); }
inline bool same_or_next(double a, double b) { return b <= up(a); }

struct Distinct {
  std::vector<double> v;
  std::vector<int> c;
};

// binning.rs:117-212
std::vector<double> greedy(const double* v, const int* c, int nd, int max_bin, int64_t total, int min_in_bin) {
  std::vector<double> bounds;
  if (nd == 0) return {INFINITY};
  if (nd <= max_bin) {
    int acc = 0;
    for (int i = 0; i + 1 < nd; ++i) {
      acc += c[i];
      if (acc < min_in_bin) continue;
      double b = up((v[i] + v[i + 1]) / 2.0);
      if (bounds.empty() || !same_or_next(bounds.back(), b)) {
        bounds.push_back(b);
        acc = 0;
      }
    }
    bounds.push_back(INFINITY);
    return bounds;
  }
  if (min_in_bin > 0) max_bin = std::max(1, std::min(max_bin, (int)(total / min_in_bin)));
  double mean = (double)total / (double)max_bin;
  int rest_bins = max_bin;
  int rest_cnt = (int)total;
  std::vector<uint8_t> big(nd);
  for (int i = 0; i < nd; ++i) {
    big[i] = (double)c[i] >= mean;
    if (big[i]) { rest_bins--; rest_cnt -= c[i]; }
  }
  mean = (double)rest_cnt / (double)rest_bins;
  std::vector<double> hi(max_bin, INFINITY), lo(max_bin, INFINITY);
  int nb = 0;
  lo[0] = v[0];
  int acc = 0;
  for (int i = 0; i + 1 < nd; ++i) {
    if (!big[i]) rest_cnt -= c[i];
    acc += c[i];
    double half = std::max(mean * (double)0.5f, 1.0);
    bool cut = big[i] || (double)acc >= mean || (big[i + 1] && (double)acc >= half);
    if (!cut) continue;
    hi[nb] = v[i];
    nb++;
    lo[nb] = v[i + 1];
    if (nb >= max_bin - 1) break;
    acc = 0;
    if (!big[i]) {
      rest_bins--;
      mean = (double)rest_cnt / (double)rest_bins;
    }
  }
  nb++;
  for (int i = 0; i + 1 < nb; ++i) {
    double b = up((hi[i] + lo[i + 1]) / 2.0);
    if (bounds.empty() || !same_or_next(bounds.back(), b)) bounds.push_back(b);
  }
  bounds.push_back(INFINITY);
  return bounds;
}

// binning.rs:215-276
std::vector<double> zero_as_one_bin(const Distinct& d, int max_bin, int64_t total, int min_in_bin) {
  const int n = (int)d.v.size();
  int neg_cnt = 0, zero_cnt = 0, pos_cnt = 0;
  for (int i = 0; i < n; ++i) {
    if (d.v[i] <= -kZero) neg_cnt += d.c[i];
    else if (d.v[i] > kZero) pos_cnt += d.c[i];
    else zero_cnt += d.c[i];
  }
  int n_neg = n;
  for (int i = 0; i < n; ++i) if (d.v[i] > -kZero) { n_neg = i; break; }
  std::vector<double> bounds;
  if (n_neg > 0 && max_bin > 1) {
    int neg_bins = (int)((double)neg_cnt / (double)(total - zero_cnt) * (double)(max_bin - 1));
    bounds = greedy(d.v.data(), d.c.data(), n_neg, std::max(neg_bins, 1), neg_cnt, min_in_bin);
    bounds.back() = -kZero;
  }
  int pos_start = -1;
  for (int i = n_neg; i < n; ++i) if (d.v[i] > kZero) { pos_start = i; break; }
  int pos_bins = max_bin - 1 - (int)bounds.size();
  if (pos_start >= 0 && pos_bins > 0) {
    std::vector<double> pb = greedy(d.v.data() + pos_start, d.c.data() + pos_start, n - pos_start, pos_bins, pos_cnt,
                                    min_in_bin);
    bounds.push_back(kZero);
    bounds.insert(bounds.end(), pb.begin(), pb.end());
  } else {
    bounds.push_back(INFINITY);
  }
  return bounds;
}

// binning.rs:94-114
bool filtered(const std::vector<int>& cnt, int total, int filter_cnt, bool numerical) {
  if (!numerical && cnt.size() > 2) return false;
  int left = 0;
  for (size_t i = 0; i + 1 < cnt.size(); ++i) {
    left = numerical ? left + cnt[i] : cnt[i];
    if (left >= filter_cnt && total - left >= filter_cnt) return false;
  }
  return true;
}

void finish(Mapper& m, const std::vector<int>& cnt, int64_t total, const lgbmhip_bin_params& p, int filter_cnt) {
  m.is_trivial = m.num_bin <= 1;
  if (!m.is_trivial && p.feature_pre_filter && filtered(cnt, (int)total, filter_cnt, m.bin_type == 0))
    m.is_trivial = true;
  if (m.is_trivial) { m.sparse_rate = 1.0; return; }
  m.default_bin = m.to_bin(0.0);
  int arg = 0;
  for (int i = 1; i < (int)cnt.size(); ++i) if (cnt[i] > cnt[arg]) arg = i;
  m.most_freq_bin = arg;
  double rate = (double)cnt[arg] / (double)total;
  if (m.most_freq_bin != m.default_bin && rate < kSparse) m.most_freq_bin = m.default_bin;
  m.sparse_rate = (double)cnt[m.most_freq_bin] / (double)total;
}

// binning.rs:398-573
Mapper build_mapper(std::vector<double> vals, int64_t total, bool categorical, const lgbmhip_bin_params& p,
                    int filter_cnt) {
  const int n_in = (int)vals.size();
  vals.erase(std::remove_if(vals.begin(), vals.end(), [](double x) { return std::isnan(x); }), vals.end());
  const int nv = (int)vals.size();
  int na = 0;
  int mt;
  if (!p.use_missing) mt = 0;
  else if (p.zero_as_missing) mt = 1;
  else if (nv == n_in) mt = 0;
  else { na = n_in - nv; mt = 2; }
  const int zeros = (int)(total - nv - na);
  std::sort(vals.begin(), vals.end());
  Distinct d;
  auto push = [&](double v, int c) { d.v.push_back(v); d.c.push_back(c); };
  if (nv == 0 || (vals[0] > 0.0 && zeros > 0)) push(0.0, zeros);
  if (nv > 0) push(vals[0], 1);
  for (int i = 1; i < nv; ++i) {
    if (same_or_next(vals[i - 1], vals[i])) {
      d.v.back() = vals[i];
      d.c.back()++;
      continue;
    }
    if (vals[i - 1] < 0.0 && vals[i] > 0.0) push(0.0, zeros);
    push(vals[i], 1);
  }
  if (nv > 0 && vals[nv - 1] < 0.0 && zeros > 0) push(0.0, zeros);

  Mapper m;
  m.missing_type = mt;
  m.bin_type = categorical ? 1 : 0;
  m.min_val = d.v.front();
  m.max_val = d.v.back();
  if (categorical) {
    m.bounds.clear();
    std::vector<std::pair<int, int>> cats;  // (count, category) in ascending category order
    for (size_t i = 0; i < d.v.size(); ++i) {
      int cat = (int)d.v[i];
      if (cat < 0) { na += d.c[i]; continue; }
      if (!cats.empty() && cats.back().second == cat) cats.back().first += d.c[i];
      else cats.push_back({d.c[i], cat});
    }
    std::vector<int> cnt;
    if (total - na > 0) {
      std::stable_sort(cats.begin(), cats.end(), [](const std::pair<int, int>& a, const std::pair<int, int>& b) {
        return a.first > b.first;
      });
      int cut = (int)((double)((float)(total - na) * 0.99f) + (double)0.5f);
      int distinct = (int)cats.size() + (na > 0 ? 1 : 0);
      int limit = std::min(distinct, p.max_bin);
      m.categories.push_back(-1);
      m.cat_to_bin[-1] = 0;
      cnt.push_back(0);
      m.num_bin = 1;
      int used = 0;
      size_t k = 0;
      for (; k < cats.size() && (used < cut || m.num_bin < limit); ++k) {
        if (cats[k].first < p.min_data_in_bin && k > 1) break;
        m.categories.push_back(cats[k].second);
        m.cat_to_bin[cats[k].second] = m.num_bin;
        used += cats[k].first;
        cnt.push_back(cats[k].first);
        m.num_bin++;
      }
      m.missing_type = (k == cats.size() && na == 0) ? 0 : 2;
      cnt[0] = (int)(total - used);
    }
    finish(m, cnt, total, p, filter_cnt);
    return m;
  }
  std::vector<double> b;
  if (mt == 2) {
    b = zero_as_one_bin(d, p.max_bin - 1, total - na, p.min_data_in_bin);
    b.push_back(NAN);
  } else {
    b = zero_as_one_bin(d, p.max_bin, total, p.min_data_in_bin);
    if (mt == 1 && b.size() == 2) mt = 0;
  }
  m.missing_type = mt;
  m.num_bin = (int)b.size();
  m.bounds = b;
  std::vector<int> cnt(m.num_bin, 0);
  int ib = 0;
  for (size_t i = 0; i < d.v.size(); ++i) {
    while (d.v[i] > b[ib] && ib < m.num_bin - 1) ib++;
    cnt[ib] += d.c[i];
  }
  if (mt == 2) cnt[m.num_bin - 1] = na;
  finish(m, cnt, total, p, filter_cnt);
  return m;
}

template <class F>
void parallel_for(int64_t n, int threads, F&& fn) {
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  threads = (int)std::min<int64_t>(threads, std::max<int64_t>(n, 1));
  if (threads <= 1) {
    for (int64_t i = 0; i < n; ++i) fn(i);
    return;
  }
  std::atomic<int64_t> next{0};
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&] {
      for (int64_t i; (i = next.fetch_add(1)) < n;) fn(i);
    });
  for (auto& th : pool) th.join();
}

}  // namespace

// binning.rs:637-667
uint32_t Mapper::to_bin(double value) const {
  if (bin_type == 1) {
    if (std::isnan(value)) return 0;
    int v = sat_i32(value);
    if (v < 0) return 0;
    auto it = cat_to_bin.find(v);
    return it == cat_to_bin.end() ? 0u : (uint32_t)it->second;
  }
  if (std::isnan(value)) {
    if (missing_type == 2) return (uint32_t)(num_bin - 1);
    value = 0.0;
  }
  int lo = 0, hi = num_bin - 1 - (missing_type == 2 ? 1 : 0);
  while (lo < hi) {
    int mid = (hi + lo - 1) / 2;
    if (value <= bounds[mid]) hi = mid; else lo = mid + 1;
  }
  return (uint32_t)lo;
}

// The number of feature groups the reference's exclusive feature bundling
// forms over the used features (feature_groups.rs:20-211, upstream_groups
// with enable_bundle and is_enable_sparse on). This engine stores every
// feature on its own; the reference's bagging reads this count to choose its
// subset mode (sample_strategy.rs:198-201). nz[f]: the sample positions at
// which used feature f holds a kept value (|v| > kZeroThreshold or NaN),
// ascending, with the values in vals[f].
namespace {

// feature_groups.rs:21-42: a feature whose zero bin is not its most frequent
// bin occupies every sample position outside its most frequent bin
std::vector<int> occupied(const Mapper& m, const std::vector<int>& nz, const std::vector<double>& vals, int total) {
  if (m.default_bin == m.most_freq_bin) return nz;
  std::vector<int> out;
  size_t j = 0;
  for (int i = 0; i < total; ++i) {
    const bool kept = j < nz.size() && nz[j] == i;
    if (!kept || m.to_bin(vals[j]) != m.most_freq_bin) out.push_back(i);
    if (kept) ++j;
  }
  return out.empty() ? nz : out;
}

// feature_groups.rs:62-158 find_groups: groups after the greedy pass and the
// merge of the sparse ones into one trailing group
int greedy_group_count(const std::vector<int>& order, const std::vector<std::vector<int>>& occ, int64_t total,
                       int64_t n) {
  const int64_t max_conflict = total / 10000;
  Lcg rng((int)n);
  std::vector<std::vector<char>> mark;
  std::vector<int64_t> data_cnt, used_cnt;
  for (int f : order) {
    const std::vector<int>& ix = occ[f];
    const int64_t cur = (int64_t)ix.size();
    std::vector<int> avail;
    for (int g = 0; g < (int)mark.size(); ++g)
      if (data_cnt[g] + cur <= total + max_conflict) avail.push_back(g);
    std::vector<int> search;
    if (!avail.empty()) {
      const int last = (int)avail.size() - 1;
      search.push_back(avail.back());
      for (int k : rng.sample(last, std::min(last, 99))) search.push_back(avail[k]);
    }
    int pick = -1;
    int64_t pick_cnt = 0;
    for (int g : search) {
      const int64_t room = max_conflict - data_cnt[g] + used_cnt[g];
      int64_t cnt = 0;
      bool over = false;
      for (int i : ix) {
        cnt += mark[g][i];
        if (cnt > room) {
          over = true;
          break;
        }
      }
      if (!over && cnt <= room && cnt <= cur / 2) {
        pick = g;
        pick_cnt = cnt;
        break;
      }
    }
    if (pick < 0) {
      pick = (int)mark.size();
      mark.emplace_back((size_t)total, 0);
      data_cnt.push_back(0);
      used_cnt.push_back(0);
      pick_cnt = 0;
    }
    data_cnt[pick] += cur;
    used_cnt[pick] += cur - pick_cnt;
    for (int i : ix) mark[pick][i] = 1;
  }
  int dense = 0;
  bool sparse = false;
  for (size_t g = 0; g < mark.size(); ++g) {
    if ((double)used_cnt[g] / (double)total >= 0.4) ++dense;
    else sparse = true;
  }
  return dense + (sparse ? 1 : 0);
}

int reference_group_count(const Binned& b, const std::vector<std::vector<int>>& nz,
                          const std::vector<std::vector<double>>& vals, int64_t total) {
  const int F = (int)b.used.size();
  if (F == 0 || total == 0) return F;
  std::vector<std::vector<int>> occ(F);
  for (int f = 0; f < F; ++f) occ[f] = occupied(b.mappers[b.used[f]], nz[f], vals[f], (int)total);
  std::vector<int> by_col(F), by_cnt(F);
  for (int f = 0; f < F; ++f) by_col[f] = by_cnt[f] = f;
  // feature_groups.rs:181-184: the densest columns first (by kept values), stable
  std::stable_sort(by_cnt.begin(), by_cnt.end(), [&](int x, int y) { return nz[x].size() > nz[y].size(); });
  return std::min(greedy_group_count(by_col, occ, total, b.n), greedy_group_count(by_cnt, occ, total, b.n));
}

}  // namespace

Binned* find_mappers(const float* X, int64_t n, int ncol, const int32_t* cats, int n_cat, const lgbmhip_bin_params& p) {
  if (n <= 0 || ncol <= 0) throw Error(LGBMHIP_INVALID_DATA, "dataset has no rows or no feature columns");
  if (n > INT32_MAX) throw Error(LGBMHIP_INVALID_DATA, "more than 2^31-1 rows");
  std::unique_ptr<Binned> out(new Binned());
  out->n = n;
  out->ncol = ncol;
  std::vector<uint8_t> is_cat(ncol, 0);
  for (int i = 0; i < n_cat; ++i) if (cats[i] >= 0 && cats[i] < ncol) is_cat[cats[i]] = 1;
  Lcg rng(p.data_random_seed);
  std::vector<int> sample = rng.sample((int)n, (int)std::min<int64_t>(n, std::max(p.bin_construct_sample_cnt, 0)));
  const int64_t total = (int64_t)sample.size();
  const int filter_cnt = (int)((double)p.min_data_in_leaf * (double)total / (double)n);
  out->mappers.resize(ncol);
  parallel_for(ncol, p.num_threads, [&](int64_t c) {
    std::vector<double> vals;
    vals.reserve(total);
    for (int r : sample) {
      double v = (double)X[(size_t)r * ncol + c];
      if (std::fabs(v) > kZero || std::isnan(v)) vals.push_back(v);
    }
    out->mappers[c] = build_mapper(std::move(vals), total, is_cat[c] != 0, p, filter_cnt);
  });
  for (int c = 0; c < ncol; ++c) {
    if (out->mappers[c].is_trivial) continue;
    if (out->mappers[c].num_bin > 256)
      throw Error(LGBMHIP_UNSUPPORTED, "a feature has more than 256 bins (u8 bin storage)");
    out->used.push_back(c);
  }
  // the sampled kept values of the used features (matrix.rs:198-214)
  const int F = (int)out->used.size();
  std::vector<std::vector<int>> nz(F);
  std::vector<std::vector<double>> vals(F);
  parallel_for(F, p.num_threads, [&](int64_t f) {
    const int c = out->used[f];
    for (int64_t i = 0; i < total; ++i) {
      const double v = (double)X[(size_t)sample[i] * ncol + c];
      if (std::fabs(v) > kZero || std::isnan(v)) {
        nz[f].push_back((int)i);
        vals[f].push_back(v);
      }
    }
  });
  out->num_groups = reference_group_count(*out, nz, vals, total);
  return out.release();
}

Binned* bin_dense(const float* X, int64_t n, int ncol, const int32_t* cats, int n_cat, const lgbmhip_bin_params& p) {
  std::unique_ptr<Binned> out(find_mappers(X, n, ncol, cats, n_cat, p));
  const int F = (int)out->used.size();
  out->bins.resize((size_t)n * F);
  const int64_t chunk = 16384;
  parallel_for((n + chunk - 1) / chunk, p.num_threads, [&](int64_t ch) {
    int64_t r0 = ch * chunk, r1 = std::min(n, r0 + chunk);
    for (int64_t r = r0; r < r1; ++r)
      for (int f = 0; f < F; ++f)
        out->bins[(size_t)r * F + f] = (uint8_t)out->mappers[out->used[f]].to_bin((double)X[(size_t)r * ncol + out->used[f]]);
  });
  return out.release();
}

void apply_bins(const Binned& b, const float* X, int64_t n, int ncol, uint8_t* out) {
  if (ncol != b.ncol) throw Error(LGBMHIP_INVALID_DATA, "column count differs from the binned training data");
  const int F = (int)b.used.size();
  const int64_t chunk = 16384;
  parallel_for((n + chunk - 1) / chunk, 0, [&](int64_t ch) {
    int64_t r0 = ch * chunk, r1 = std::min(n, r0 + chunk);
    for (int64_t r = r0; r < r1; ++r)
      for (int f = 0; f < F; ++f)
        out[(size_t)r * F + f] = (uint8_t)b.mappers[b.used[f]].to_bin((double)X[(size_t)r * ncol + b.used[f]]);
  });
}

}  // namespace lgbmhip
