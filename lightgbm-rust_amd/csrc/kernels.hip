// CDNA4 (gfx950) kernels of the GBDT hot path.
//
// Tree growth runs as a device-resident state machine (StepDev): every
// kernel reads the leaf it works on, its row segment and the split to apply
// from device memory, so one tree is a fixed sequence of launches with no
// host round trip (and can be captured in a hipGraph):
//
//   tree_init -> hist -> split(root) -> { part_count -> part_scatter -> hist -> split } x (num_leaves-1)
//             -> [renew] -> leaf_values -> score
//
// Determinism: FP64 histograms are accumulated as exact 64-bit fixed-point
// integers (pre-rounding each f32 gradient to a power-of-two grid sized from
// the leaf's row count and the iteration's max |g|, after Demmel & Nguyen's
// reproducible-summation idea), so per-bin sums are independent of block
// count, tiling and atomic order; the quantized form sums packed int8
// gradients exactly. Threshold-search prefix sums run in the reference's
// sequential order (FP64, one chain per wave), so identical histograms give
// identical splits.

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>

#include "engine.h"

namespace lgbmhip {

#define KEPS ((double)1e-15f)
#define KZERO ((double)1e-35f)

__device__ __forceinline__ double round0(double v) { return (v >= -KZERO && v <= KZERO) ? 0.0 : v; }
__device__ __forceinline__ int round_int(double x) { return (int)(x + (double)0.5f); }
__device__ __forceinline__ double sgn(double x) { return (double)((x > 0.0) - (x < 0.0)); }
__device__ __forceinline__ double thr_l1(double s, double l1) { return sgn(s) * fmax(fabs(s) - l1, 0.0); }

// split.rs:137-161
__device__ __forceinline__ double leaf_output_t(double g, double h, const SplitParamsDev& p, bool use_l1, bool use_max,
                                bool use_smooth, int nd, double po) {
  double ret = use_l1 ? -thr_l1(g, p.lambda_l1) / (h + p.lambda_l2) : -g / (h + p.lambda_l2);
  if (use_max && p.max_delta_step > 0.0 && fabs(ret) > p.max_delta_step) ret = sgn(ret) * p.max_delta_step;
  if (use_smooth) {
    double n = (double)nd;
    ret = ret * (n / p.path_smooth) / (n / p.path_smooth + 1.0) + po / (n / p.path_smooth + 1.0);
  }
  return ret;
}
__device__ __forceinline__ double leaf_output(double g, double h, const SplitParamsDev& p, int nd, double po) {
  return leaf_output_t(g, h, p, p.lambda_l1 > 0.0, p.max_delta_step > 0.0, p.path_smooth > KEPS, nd, po);
}
// split.rs:183-201
__device__ __forceinline__ double leaf_gain(double g, double h, const SplitParamsDev& p, int nd, double po) {
  bool l1 = p.lambda_l1 > 0.0;
  if (!(p.max_delta_step > 0.0) && !(p.path_smooth > KEPS)) {
    if (l1) {
      double s = thr_l1(g, p.lambda_l1);
      return (s * s) / (h + p.lambda_l2);
    }
    return (g * g) / (h + p.lambda_l2);
  }
  double out = leaf_output(g, h, p, nd, po);
  double s = l1 ? thr_l1(g, p.lambda_l1) : g;
  return -(2.0 * s * out + (h + p.lambda_l2) * out * out);
}
__device__ __forceinline__ double split_gain(double lg, double lh, double rg, double rh, const SplitParamsDev& p,
                                             int lc, int rc, double po) {
  return leaf_gain(lg, lh, p, lc, po) + leaf_gain(rg, rh, p, rc, po);
}

// Fixed-point exponent for summing `count` values of magnitude < 2^e (e from
// the max's float bits) into int64 without overflow: |sum * 2^s| < 2^62.
__device__ __host__ __forceinline__ int fix_shift(long long count, unsigned max_bits) {
  int r = 1;
  while ((1LL << r) <= count && r < 62) ++r;   // 2^r > count
  float m;
  memcpy(&m, &max_bits, 4);
  int e = 0;
  if (m > 0.0f) {
    int ex;
    frexpf(m, &ex);  // m = f * 2^ex, f in [0.5, 1): m < 2^ex
    e = ex;
  }
  return 62 - r - e;
}
__device__ __forceinline__ long long to_fixed(float v, int s) { return __double2ll_rn(ldexp((double)v, s)); }

__device__ __forceinline__ int8_t to_i8(double x) {
  if (isnan(x) || !(x >= -2147483648.0 && x < 2147483648.0)) return 0;
  return (int8_t)(int32_t)x;
}

// A split is better: higher gain, then smaller feature index (split.rs:114-121).
__device__ __forceinline__ bool better(double ga, int fa, double gb, int fb) {
  if (ga != gb) return ga > gb;
  int a = fa < 0 ? INT32_MAX : fa, b = fb < 0 ? INT32_MAX : fb;
  return a < b;
}

__device__ __forceinline__ void reset_rec(SplitRec& r) {
  r.gain = -INFINITY;
  r.feature = -1;
  r.threshold = 0;
  r.default_left = 1;
  r.left_output = r.right_output = 0.0;
  r.lsg = r.lsh = r.rsg = r.rsh = 0.0;
  r.lgh = r.rgh = 0;
  r.left_count = r.right_count = 0;
  r.num_cat = 0;
  r.leaf = -1;
  for (int i = 0; i < kCatWords; ++i) r.cat_bits[i] = 0;
}

// ---------------------------------------------------------------- block helpers
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ unsigned wave_max_u(unsigned v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, (unsigned)__shfl_xor((int)v, o, 64));
  return v;
}

// Block-wide reductions for 256-thread blocks, result in thread 0: one
// global atomic per block instead of one per wave (same-address atomics
// serialize in the memory system).
__device__ __forceinline__ unsigned block_max_u(unsigned v) {
  __shared__ unsigned sm_max[4];
  v = wave_max_u(v);
  if ((threadIdx.x & 63) == 0) sm_max[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) v = max(max(sm_max[0], sm_max[1]), max(sm_max[2], sm_max[3]));
  __syncthreads();
  return v;
}
__device__ __forceinline__ long long block_sum_ll(long long v) {
  __shared__ long long sm_sum[4];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) sm_sum[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) v = sm_sum[0] + sm_sum[1] + sm_sum[2] + sm_sum[3];
  __syncthreads();
  return v;
}

// Orders LDS traffic between the lanes of one wave (a wave-scope barrier;
// the two waves of a search block never wait for each other).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wave reductions and scans on DPP row operations (no LDS round trips).
// Minimum of an int over the wave, in every lane.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ int dpp_min_step(int v) {
  return min(v, __builtin_amdgcn_update_dpp(v, v, CTRL, ROWMASK, 0xf, false));
}
__device__ __forceinline__ int wave_min_i32(int v) {
  v = dpp_min_step<0xB1, 0xF>(v);    // quad_perm [1,0,3,2]
  v = dpp_min_step<0x4E, 0xF>(v);    // quad_perm [2,3,0,1]
  v = dpp_min_step<0x141, 0xF>(v);   // row_half_mirror
  v = dpp_min_step<0x140, 0xF>(v);   // row_mirror
  v = dpp_min_step<0x142, 0xA>(v);   // row_bcast:15 into rows 1, 3
  v = dpp_min_step<0x143, 0xC>(v);   // row_bcast:31 into rows 2, 3
  return __builtin_amdgcn_readlane(v, 63);
}

// The value of lane - k of the same 16-lane row (ROWMASK rows), 0 elsewhere.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ unsigned long long dpp_u64_or0(unsigned long long v) {
  const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)v, CTRL, ROWMASK, 0xf, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(v >> 32), CTRL, ROWMASK, 0xf, false);
  return ((unsigned long long)hi << 32) | lo;
}
// Inclusive prefix sum over the wave's lanes (wrapping 64-bit).
__device__ __forceinline__ unsigned long long wave_incl_scan_u64(unsigned long long v) {
  v += dpp_u64_or0<0x111, 0xF>(v);   // row_shr:1
  v += dpp_u64_or0<0x112, 0xF>(v);   // row_shr:2
  v += dpp_u64_or0<0x114, 0xF>(v);   // row_shr:4
  v += dpp_u64_or0<0x118, 0xF>(v);   // row_shr:8
  v += dpp_u64_or0<0x142, 0xA>(v);   // row_bcast:15 into rows 1, 3
  v += dpp_u64_or0<0x143, 0xC>(v);   // row_bcast:31 into rows 2, 3
  return v;
}

// The same order as packed keys: k1 = the gain as an order-preserving 64-bit
// integer (larger is better), k2 = (feature, position) with a missing
// feature last (smaller is better). Positions < 65536, features < 65535.
struct KeyRec {
  unsigned long long k1;
  unsigned k2;
  int spec;   // >= 0: split ahead of order
};

__device__ __forceinline__ unsigned long long gain_key(double g) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(g);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double key_gain(unsigned long long k) {
  const unsigned long long b = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
  return __longlong_as_double((long long)b);
}
__device__ __forceinline__ unsigned tie_key(int feat, int pos) {
  return ((unsigned)(feat < 0 ? 0xFFFF : feat) << 16) | (unsigned)pos;
}
__device__ __forceinline__ bool key_better(unsigned long long a1, unsigned a2, unsigned long long b1, unsigned b2) {
  return a1 > b1 || (a1 == b1 && a2 < b2);
}

// DPP step of a wave reduction: this lane's value and the one DPP control
// CTRL brings from another lane (rows outside ROWMASK keep their own).
template <int CTRL, int ROWMASK>
__device__ __forceinline__ unsigned dpp_from(unsigned v) {
  return (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, CTRL, ROWMASK, 0xf, false);
}

template <int CTRL, int ROWMASK>
__device__ __forceinline__ void dpp_best_step(unsigned& h, unsigned& l, unsigned& t) {
  const unsigned oh = dpp_from<CTRL, ROWMASK>(h), ol = dpp_from<CTRL, ROWMASK>(l), ot = dpp_from<CTRL, ROWMASK>(t);
  const unsigned long long a = ((unsigned long long)h << 32) | l, b = ((unsigned long long)oh << 32) | ol;
  const bool take = key_better(b, ot, a, t);
  h = take ? oh : h;
  l = take ? ol : l;
  t = take ? ot : t;
}

// Wave ArgMax of (k1, k2) pairs with DPP row operations (quad swaps, row
// mirrors, row broadcasts): lane 63 ends with the best pair, read back into
// every lane.
__device__ __forceinline__ void wave_best_key(unsigned long long& k1, unsigned& k2) {
  unsigned h = (unsigned)(k1 >> 32), l = (unsigned)k1, t = k2;
  dpp_best_step<0xB1, 0xF>(h, l, t);    // quad_perm [1,0,3,2]
  dpp_best_step<0x4E, 0xF>(h, l, t);    // quad_perm [2,3,0,1]
  dpp_best_step<0x141, 0xF>(h, l, t);   // row_half_mirror
  dpp_best_step<0x140, 0xF>(h, l, t);   // row_mirror
  dpp_best_step<0x142, 0xA>(h, l, t);   // row_bcast:15 into rows 1, 3
  dpp_best_step<0x143, 0xC>(h, l, t);   // row_bcast:31 into rows 2, 3
  h = __builtin_amdgcn_readlane(h, 63);
  l = __builtin_amdgcn_readlane(l, 63);
  t = __builtin_amdgcn_readlane(t, 63);
  k1 = ((unsigned long long)h << 32) | l;
  k2 = t;
}

// The first maximum of a gain over the wave (the smallest k among equal
// gains) in every lane; a lane without a candidate holds (-inf, INT32_MAX).
__device__ __forceinline__ void wave_first_max(double& g, int& k) {
  unsigned long long k1 = gain_key(g == 0.0 ? 0.0 : g);
  unsigned k2 = (unsigned)k;
  wave_best_key(k1, k2);
  g = key_gain(k1);
  k = (int)k2;
}


// Last block of the grid to arrive (release/acquire around a device counter).
__device__ bool last_block(int* counter) {
  __shared__ int ticket;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    ticket = atomicAdd(counter, 1);
  }
  __syncthreads();
  bool last = ticket == (int)(gridDim.x * gridDim.y) - 1;
  if (last) __threadfence();
  return last;
}

// Advances the quantizer's call counter (one start offset per discretize call).
__global__ void k_bump(int* p, int by) { *p += by; }

// Clears the per-iteration reductions of K classes (a fill kernel, so the
// iteration can be captured in a graph).
__global__ void k_iter_reset(IterDev* it, int K) {
  int k = threadIdx.x;
  if (k < K) {
    it[k].max_abs_g_bits = 0;
    it[k].max_abs_h_bits = 0;
    it[k].root_g_fix = 0;
    it[k].root_h_fix = 0;
    it[k].root_int_sum = 0;
    it[k].root_gq_sum = 0;
    it[k].root_hq_sum = 0;
    it[k].qmax_g = 0;
    it[k].qmax_h = 0;
    it[k].full_bag = 0;
    it[k].bag_sg_fix = 0;
    it[k].bag_sh_fix = 0;
    it[k].bag_isum = 0;
  }
}

// ---------------------------------------------------------------- objectives
// objective/binary.rs:101-127. One row per thread; max |g|, |h| by block
// reduction then one atomicMax per block (order-free).
__global__ void __launch_bounds__(256) k_binary_grad(const double* __restrict__ score, const float* __restrict__ label,
                                                     const float* __restrict__ weight, long long n, double sigmoid,
                                                     double lw_neg, double lw_pos, float* __restrict__ g,
                                                     float* __restrict__ h, IterDev* it) {
  constexpr int U = 4;   // rows in flight per thread
  unsigned mg = 0, mh = 0;
  const long long step = (long long)gridDim.x * blockDim.x * U;
  for (long long base = (long long)blockIdx.x * blockDim.x * U + threadIdx.x; base < n; base += step) {
    float lv[U], wv[U];
    double sv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long i = base + (long long)u * blockDim.x;
      lv[u] = 0.f;
      sv[u] = 0.0;
      wv[u] = 1.f;
      if (i < n) {
        lv[u] = label[i];
        sv[u] = score[i];
        if (weight) wv[u] = weight[i];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long i = base + (long long)u * blockDim.x;
      if (i >= n) continue;
      const bool pos = lv[u] > 0.0f;
      const double lab = pos ? 1.0 : -1.0;
      const double lw = pos ? lw_pos : lw_neg;
      const double resp = -lab * sigmoid / ((double)1.0f + exp(lab * sigmoid * sv[u]));
      const double ar = fabs(resp);
      float gv, hv;
      if (weight) {
        const double w = (double)wv[u];
        gv = (float)(resp * lw * w);
        hv = (float)(ar * (sigmoid - ar) * lw * w);
      } else {
        gv = (float)(resp * lw);
        hv = (float)(ar * (sigmoid - ar) * lw);
      }
      g[i] = gv;
      h[i] = hv;
      mg = max(mg, __float_as_uint(fabsf(gv)));
      mh = max(mh, __float_as_uint(fabsf(hv)));
    }
  }
  mg = block_max_u(mg);
  mh = block_max_u(mh);
  if (threadIdx.x == 0) {
    atomicMax(&it->max_abs_g_bits, mg);
    atomicMax(&it->max_abs_h_bits, mh);
  }
}

// objective/multiclass.rs:16-30, 126-154 (K <= 16 classes). The softmax
// terms exp(s_c - max) are recomputed per pass rather than held in a per-row
// array (which would live in scratch); exp is deterministic, so the sum and
// the probabilities are the ones a stored array would give. The per-class
// |g|, |h| maxima are reduced per wave before one LDS atomic per class.
__global__ void __launch_bounds__(256) k_multiclass_grad(const double* __restrict__ score, const float* __restrict__ label,
                                  const float* __restrict__ weight, long long n, int K, float* __restrict__ g,
                                  float* __restrict__ h, IterDev* it) {
  const double factor = (double)K / (double)((float)K - 1.0f);
  __shared__ unsigned smg[16], smh[16];
  if (threadIdx.x < 16) { smg[threadIdx.x] = 0; smh[threadIdx.x] = 0; }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const long long stride = (long long)gridDim.x * blockDim.x;
  // block-uniform trip count: every lane takes part in the wave reductions
  for (long long base = (long long)blockIdx.x * blockDim.x; base < n; base += stride) {
    const long long i = base + threadIdx.x;
    const bool ok = i < n;
    const long long r = ok ? i : 0;
    double wmax = score[r];
    for (int c = 1; c < K; ++c) {
      const double v = score[(long long)c * n + r];
      wmax = v < wmax ? wmax : v;
    }
    double wsum = 0.0;
    for (int c = 0; c < K; ++c) wsum += exp(score[(long long)c * n + r] - wmax);
    const int li = (int)label[r];
    const double w = weight ? (double)weight[r] : 1.0;
    for (int c = 0; c < K; ++c) {
      const double pc = exp(score[(long long)c * n + r] - wmax) / wsum;
      double gg = li == c ? pc - (double)1.0f : pc;
      double hh = factor * pc * ((double)1.0f - pc);
      if (weight) {
        gg *= w;
        hh *= w;
      }
      const float gv = (float)gg, hv = (float)hh;
      if (ok) {
        g[(long long)c * n + i] = gv;
        h[(long long)c * n + i] = hv;
      }
      unsigned mg = ok ? __float_as_uint(fabsf(gv)) : 0u, mh = ok ? __float_as_uint(fabsf(hv)) : 0u;
      for (int o = 32; o > 0; o >>= 1) {
        mg = max(mg, (unsigned)__shfl_xor((int)mg, o, 64));
        mh = max(mh, (unsigned)__shfl_xor((int)mh, o, 64));
      }
      if (lane == 0) {
        atomicMax(&smg[c], mg);
        atomicMax(&smh[c], mh);
      }
    }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < K; c += blockDim.x) {
    atomicMax(&it[c].max_abs_g_bits, smg[c]);
    atomicMax(&it[c].max_abs_h_bits, smh[c]);
  }
}

// objective/rank.rs:192-257 LambdarankNdcg::one_query, one block per query.
// The reference accumulates each document's f32 lambda/hessian in (i, j)
// pair-loop order and the normalizer sum_lambdas as one f64 chain in the
// same order; k_lambdarank keeps every one of those orders (below).
struct RankArgs {
  const double* score;
  const float* label;
  const int* qb;
  int Q;
  const double* inv_max_dcg;
  const double* table;       // 1M-entry sigmoid table (rank.rs:166-179)
  double min_in, max_in, idx_factor, sigmoid;
  int norm, trunc;
  const double* label_gain;  // 31 entries
  const double* discount;    // 10000 entries
  const int* order;          // [Q]: the queries, longest first (block b runs query order[b])
  int* rank;                 // scratch [N]: rank of each row within its query
  double* by_rank_s;         // scratch [N]: score at each rank
  float* by_rank_l;          // scratch [N]: label at each rank
  float* lam;                // scratch [N]: accumulators by rank
  float* hes;
  float* g;
  float* h;
};

constexpr int kRankThreads = 256;

__device__ __forceinline__ double rank_sigmoid(const RankArgs& a, double s) {
  if (s <= a.min_in) return a.table[0];
  if (s >= a.max_in) return a.table[1024 * 1024 - 1];
  // in range the product lies in (0, 2^20]: a 32-bit truncation is the
  // reference's `as usize`; the clamp only bounds the rounding case 2^20
  const unsigned ix = (unsigned)((s - a.min_in) * a.idx_factor);
  return a.table[min(ix, 1024u * 1024u - 1u)];
}

// One pair's contribution (rank.rs:213-247), labels differ, from the
// anchor's and the partner's values: d = s_anchor - s_partner, their label
// gains ga, gb and rank discounts da, db. The reference takes high minus low
// throughout; with the anchor low that negates d and ga - gb, exactly (IEEE
// subtraction is sign-symmetric), and leaves |da - db| unchanged, so these
// terms are the reference's bit for bit. sig: the sigmoid of the gap. fl / fh:
// the f32 lambda and hessian terms, plo the f64 lambda before its cast (the
// normalizer's chain subtracts 2 plo).
__device__ __forceinline__ void rank_pair(const RankArgs& a, double imd, bool scale_by_delta, bool anchor_high,
                                          double d, double ga, double gb, double da, double db, double sig,
                                          float& fl, float& fh, double& plo) {
  const double gd = ga - gb;
  const double gap = anchor_high ? gd : -gd;
  const double pd = fabs(da - db);
  double dp = gap * pd * imd;
  if (scale_by_delta) dp /= (double)0.01f + fabs(d);
  double pl = sig;
  double ph = pl * (1.0 - pl);
  pl *= -a.sigmoid * dp;
  ph *= a.sigmoid * a.sigmoid * dp;
  fl = (float)pl;
  fh = (float)ph;
  plo = pl;
}

// The ordered sums of one query. Every pair term is a function of scores,
// labels and ranks only; the reference's order binds only the sums: each
// document's f32 lambda and hessian (its contributions as a partner of the
// earlier anchors, in anchor order, then, for an anchor, its own terms in
// partner order) and the f64 normalizer (every pair, anchor by anchor, in
// partner order). So anchors go in steps of G: the block computes the
// step's pair terms (a thread owns a partner and applies its contributions
// in anchor order) and stages them in LDS; then wave 0's lanes run the G
// anchors' own chains side by side while lane 0 of wave 1 runs the
// normalizer chain over the step's terms. A query of at most kRankLds
// documents keeps its arrays by rank in LDS and takes as many anchors per
// step as its staged terms allow; a longer one keeps them in the global
// scratch and takes one anchor per step, its partners in chunks.
constexpr int kRankLds = 512;
constexpr int kRankStage = 896;    // pair terms staged per step
constexpr int kRankGroup = 4;      // anchors per step, at most (more spill)
struct RankLds {
  double st_p[kRankStage + 32];   // + the zero padding the normalizer chain reads
  float st_l[kRankStage], st_h[kRankStage];
  double rs[kRankLds];
  float rl[kRankLds], lam[kRankLds], hes[kRankLds];
  double sgain[31];
  double sfac, S;
  int sdo;
};

template <bool small>
__device__ __forceinline__ void lambdarank_query(const RankArgs& a, RankLds& R, int q, int start, int cnt) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double* s = a.score + start;
  const float* lab = a.label + start;
  // arrays by rank: LDS or the global scratch
  double* rs = small ? R.rs : a.by_rank_s + start;
  float* rl = small ? R.rl : a.by_rank_l + start;
  float* lam = small ? R.lam : a.lam + start;
  float* hes = small ? R.hes : a.hes + start;
  const double* disc = a.discount;   // 80 KB, cache-resident
  int* rank = a.rank + start;
  if (tid < 31) R.sgain[tid] = a.label_gain[tid];
  const double imd = a.inv_max_dcg[q];
  // ---- stable descending order: rank = #{j : s_j > s_i or (s_j == s_i and
  // j < i)}, i.e. j < i counts when !(s_j < s_i), j > i when s_j > s_i. The
  // scores are compared from LDS, kRankStage at a time. Per pass, tpd threads
  // (a power of two, adjacent lanes) share a document, each counting every
  // tpd-th j; batches of j wholly before or after all of a wave's documents
  // take the one comparison that applies there.
  for (int i0 = 0; i0 < cnt;) {
    int tpd = 1;
    while (tpd < 64 && (cnt - i0) * tpd * 2 <= kRankThreads) tpd *= 2;
    const int part = tid & (tpd - 1), step = 8 * tpd;
    const int i = i0 + tid / tpd;
    const double si = s[i < cnt ? i : cnt - 1];
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int iw_lo = i0 + wv * 64 / tpd, iw_hi = i0 + (wv * 64 + 63) / tpd;
    int r = 0;
    for (int j0 = 0; j0 < cnt; j0 += kRankStage) {
      const int jn = min(cnt, j0 + kRankStage);
      __syncthreads();
      for (int j = j0 + tid; j < jn; j += kRankThreads) R.st_p[j - j0] = s[j];
      __syncthreads();
      const double* st = R.st_p - j0 + part;
      int jb = j0;
      for (; jb + step <= min(jn, iw_lo); jb += step) {   // every j < i
        double s8[8];
#pragma unroll
        for (int v = 0; v < 8; ++v) s8[v] = st[jb + v * tpd];
#pragma unroll
        for (int v = 0; v < 8; ++v) r += !(s8[v] < si);
      }
      for (; jb + step <= jn && jb <= iw_hi; jb += step) {
        double s8[8];
#pragma unroll
        for (int v = 0; v < 8; ++v) s8[v] = st[jb + v * tpd];
#pragma unroll
        for (int v = 0; v < 8; ++v) {
          const int jj = jb + part + v * tpd;
          r += jj < i ? !(s8[v] < si) : (s8[v] > si);
        }
      }
      for (; jb + step <= jn; jb += step) {   // every j > i
        double s8[8];
#pragma unroll
        for (int v = 0; v < 8; ++v) s8[v] = st[jb + v * tpd];
#pragma unroll
        for (int v = 0; v < 8; ++v) r += s8[v] > si;
      }
      for (int j = jb + part; j < jn; j += tpd) {
        const double sj = R.st_p[j - j0];
        r += j < i ? !(sj < si) : (sj > si);
      }
    }
    for (int o = 1; o < tpd; o <<= 1) r += __shfl_xor(r, o, 64);
    if (part == 0 && i < cnt) {
      rank[i] = r;
      rs[r] = si;
      rl[r] = lab[i];
      lam[r] = 0.0f;
      hes[r] = 0.0f;
    }
    i0 += kRankThreads / tpd;
  }
  if (tid == 0) R.S = 0.0;
  __syncthreads();
  const double best = rs[0];
  int worst_idx = cnt - 1;
  if (worst_idx > 0 && rs[worst_idx] == -INFINITY) worst_idx--;
  const double worst = rs[worst_idx];
  const bool scale_by_delta = a.norm && best != worst;
  const int imax = min(cnt - 1, max(a.trunc, 0));
  float anc_l = 0.0f, anc_h = 0.0f;   // wave 0 lane u: anchor g + u's own chains
  double S = 0.0;                     // wave 1 lane 0: the normalizer
  for (int g = 0; g < imax;) {
    // the step: anchors g .. g + G - 1, partners j0 .. j1 - 1 (W of them)
    const int G = small ? min(min(kRankGroup, imax - g), max(1, kRankStage / cnt)) : 1;
    for (int j0 = g + 1; j0 < cnt; j0 += kRankStage) {
      const int j1 = small ? cnt : min(cnt, j0 + kRankStage);
      const int W = j1 - j0;
      // the step's terms, compact in the normalizer's order: anchor u's
      // partners k0_u = max(j0, g + u + 1) .. j1 - 1 from off_u on (with
      // j0 = g + 1 unless G = 1: off_u = u W - u (u - 1) / 2), then zeros
      const int len = G * W - G * (G - 1) / 2;
      if (tid < 32) R.st_p[len + tid] = 0.0;
      // ---- pair terms: thread t owns partners j0 + t, j0 + t + 256, ...; the
      // step's anchors in registers. An anchor scored -inf is skipped (no
      // pair of it counts); its staged terms are the identities below.
      double sa[kRankGroup];
      float la[kRankGroup];
#pragma unroll
      for (int u = 0; u < kRankGroup; ++u) {
        sa[u] = u < G ? rs[g + u] : -INFINITY;
        la[u] = u < G ? rl[g + u] : 0.0f;
      }
      for (int j = j0 + tid; j < j1; j += kRankThreads) {
        const double sb = rs[j];
        const float lb = rl[j];
        const double gb = R.sgain[(int)lb], db = disc[j];
        // the sigmoid lookups first (a global table), all in flight together;
        // the gap high minus low is d with the anchor high, else -d
        double sig[kRankGroup];
#pragma unroll
        for (int u = 0; u < kRankGroup; ++u) {
          sig[u] = 0.0;
          if (j > g + u && sa[u] != -INFINITY && sb != -INFINITY && la[u] != lb) {
            const double d = sa[u] - sb;
            sig[u] = rank_sigmoid(a, la[u] > lb ? d : -d);
          }
        }
        float pl_ = lam[j], ph_ = hes[j];
#pragma unroll
        for (int u = 0; u < kRankGroup; ++u) {
          const int i = g + u;
          if (u >= G || j <= i) continue;
          const int e = u * W - u * (u - 1) / 2 + j - max(j0, i + 1);
          if (sa[u] != -INFINITY && sb != -INFINITY && la[u] != lb) {
            const bool anchor_high = la[u] > lb;
            float fl, fh;
            double p;
            rank_pair(a, imd, scale_by_delta, anchor_high, sa[u] - sb, R.sgain[(int)la[u]], gb, disc[i], db, sig[u], fl,
                      fh, p);
            // the partner gets its one contribution of anchor i
            if (anchor_high) pl_ -= fl;
            else pl_ += fl;
            ph_ += fh;
            R.st_l[e] = anchor_high ? fl : -fl;
            R.st_h[e] = fh;
            R.st_p[e] = 2.0 * p;   // exact: the normalizer subtracts 2 p
          } else {
            // a skipped pair adds nothing: x + (-0.0f) == x, and S - 0.0 == S
            // as S never holds -0.0
            R.st_l[e] = -0.0f;
            R.st_h[e] = -0.0f;
            R.st_p[e] = 0.0;
          }
        }
        lam[j] = pl_;
        hes[j] = ph_;
      }
      __syncthreads();
      // ---- the chains: the anchors' own on wave 0 (lane u: anchor g + u),
      // the normalizer on wave 1
      if (wave == 0 && lane < G) {
        const int i = g + lane;
        // the anchor's sums so far: its contributions as a partner
        if (j0 == g + 1) {
          anc_l = lam[i];
          anc_h = hes[i];
        }
        const int off = lane * W - lane * (lane - 1) / 2, n = j1 - max(j0, i + 1);
        const float* l = R.st_l + off;
        const float* h = R.st_h + off;
        int k = 0;
        for (; k + 8 <= n; k += 8) {
          float l8[8], h8[8];
#pragma unroll
          for (int v = 0; v < 8; ++v) {
            l8[v] = l[k + v];
            h8[v] = h[k + v];
          }
#pragma unroll
          for (int v = 0; v < 8; ++v) {
            anc_l += l8[v];
            anc_h += h8[v];
          }
        }
        for (; k < n; ++k) {
          anc_l += l[k];
          anc_h += h[k];
        }
        if (j1 == cnt) {
          lam[i] = anc_l;
          hes[i] = anc_h;
        }
      }
      if (wave == 1 && lane == 0) {
        // one run of len terms, zero-padded to 16: two batches of 8 in
        // turn, each reloaded while the other is added
        double x[8], y[8];
#pragma unroll
        for (int v = 0; v < 8; ++v) {
          x[v] = R.st_p[v];
          y[v] = R.st_p[8 + v];
        }
        for (int k = 0; k < len; k += 16) {
#pragma unroll
          for (int v = 0; v < 8; ++v) S -= x[v];
#pragma unroll
          for (int v = 0; v < 8; ++v) x[v] = R.st_p[k + 16 + v];
#pragma unroll
          for (int v = 0; v < 8; ++v) S -= y[v];
#pragma unroll
          for (int v = 0; v < 8; ++v) y[v] = R.st_p[k + 24 + v];
        }
      }
      __syncthreads();
    }
    g += G;
  }
  if (wave == 1 && lane == 0) {
    R.sdo = a.norm && S > 0.0;
    R.sfac = R.sdo ? log2(1.0 + S) / S : 1.0;
  }
  __syncthreads();
  const double sfac = R.sfac;
  const bool sdo = R.sdo != 0;
  for (int d = tid; d < cnt; d += kRankThreads) {
    const int r = rank[d];
    float l = lam[r], hh = hes[r];
    if (sdo) {
      l = (float)((double)l * sfac);
      hh = (float)((double)hh * sfac);
    }
    a.g[start + d] = l;
    a.h[start + d] = hh;
  }
}

// Six blocks per compute unit (at most 80 registers, 25 KB of LDS each): the
// serial chains of more queries interleave on each SIMD.
__global__ void __launch_bounds__(kRankThreads, 6) k_lambdarank(RankArgs a) {
  if ((int)blockIdx.x >= a.Q) return;
  const int q = a.order[blockIdx.x];   // longest queries first
  const int start = a.qb[q], cnt = a.qb[q + 1] - start;
  if (cnt <= 0) return;
  __shared__ RankLds R;
  if (cnt <= kRankLds) lambdarank_query<true>(a, R, q, start, cnt);
  else lambdarank_query<false>(a, R, q, start, cnt);
}

// max |g|, |h| after the ranking gradients, or over the rows of a bag
// (rows[0..n), when given) (order-free).
__global__ void __launch_bounds__(256) k_absmax_it(const float* g, const float* h, long long n, IterDev* it,
                                                   const unsigned* rows) {
  unsigned mg = 0, mh = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const long long r = rows ? (long long)rows[i] : i;
    mg = max(mg, __float_as_uint(fabsf(g[r])));
    mh = max(mh, __float_as_uint(fabsf(h[r])));
  }
  mg = block_max_u(mg);
  mh = block_max_u(mh);
  if (threadIdx.x == 0) {
    atomicMax(&it->max_abs_g_bits, mg);
    atomicMax(&it->max_abs_h_bits, mh);
  }
}

// ---------------------------------------------------------------- tree init
// Gradients of class k in buffer 0 (the root, rows in order; the identity
// row index of buffer 0 is written once per training run), fixed-point root
// sums (learner/mod.rs:580-628).
__global__ void __launch_bounds__(256) k_tree_init(long long n, const float* __restrict__ g, const float* __restrict__ h,
                            unsigned* __restrict__ idx0, float2* __restrict__ gh0, IterDev* it,
                            const unsigned* __restrict__ bag) {
  const int sg = fix_shift(n, it->max_abs_g_bits), sh = fix_shift(n, it->max_abs_h_bits);
  long long ag = 0, ah = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const long long r = bag ? (long long)bag[i] : i;   // position i of the root holds row r
    const float gv = g[r], hv = h[r];
    if (idx0) idx0[i] = (unsigned)i;
    gh0[i] = make_float2(gv, hv);
    ag += to_fixed(gv, sg);
    ah += to_fixed(hv, sh);
  }
  ag = block_sum_ll(ag);
  ah = block_sum_ll(ah);
  if (threadIdx.x == 0) {
    atomicAdd((unsigned long long*)&it->root_g_fix, (unsigned long long)ag);
    atomicAdd((unsigned long long*)&it->root_h_fix, (unsigned long long)ah);
  }
}

// Quantized form: grad/hess -> int8 by stochastic rounding (learner/quant.rs:92-129),
// packed (gq, hq) into buffer 0; exact int root sums. idx0 / gq_out / hq_out
// are written when given (the stage entry point).
__global__ void __launch_bounds__(256) k_discretize(long long n, const float* __restrict__ g, const float* __restrict__ h,
                             const double* __restrict__ grv, const double* __restrict__ hrv, int bins, int stochastic,
                             int constant_hessian, const int* __restrict__ starts, const int* call_ctr, int call_off,
                             IterDev* it,
                             unsigned* __restrict__ idx0, char2* __restrict__ q0, int8_t* __restrict__ gq_out,
                             int8_t* __restrict__ hq_out, const unsigned* __restrict__ bag) {
  const double mg = (double)__uint_as_float(it->max_abs_g_bits);
  const double mh = (double)__uint_as_float(it->max_abs_h_bits);
  const double gs = mg / (double)(bins / 2);
  const double hs = constant_hessian ? mh : mh / (double)bins;
  const double inv_g = 1.0 / gs, inv_h = 1.0 / hs;
  const long long start = starts[*call_ctr + call_off];
  const long long start_n = n > 0 ? start % n : 0;   // a bag has fewer rows than the offsets' range
  long long ai = 0, agq = 0, ahq = 0;
  unsigned mq = 0;   // max |gq| (low 16 bits) and max u8(hq) (high 16 bits), for the histogram's 32-bit bins
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    // position i of the root (of the bag: row bag[i], quant.rs:92-129 with rows)
    const long long r = bag ? (long long)bag[i] : i;
    const double gv = (double)g[r];
    double rg = 0.5, rh = 0.5;
    if (stochastic) {
      long long pos = i + start_n;   // (i + start) mod n
      if (pos >= n) pos -= n;
      rg = grv[pos];
      rh = hrv[pos];
    }
    const int8_t a = gv >= 0.0 ? to_i8(gv * inv_g + rg) : to_i8(gv * inv_g - rg);
    const int8_t b = constant_hessian ? (int8_t)1 : to_i8((double)h[r] * inv_h + rh);
    if (idx0) idx0[i] = (unsigned)i;
    q0[i] = make_char2(a, b);
    if (gq_out) {
      gq_out[i] = a;
      hq_out[i] = b;
    }
    ai += (long long)(((unsigned long long)(long long)a << 32) | (unsigned long long)(uint8_t)b);
    agq += a;
    ahq += b;
    const unsigned ag = (unsigned)(a < 0 ? -(int)a : (int)a), bh = (unsigned)(uint8_t)b;
    mq = max(mq & 0xFFFFu, ag) | (max(mq >> 16, bh) << 16);
  }
  ai = block_sum_ll(ai);
  agq = block_sum_ll(agq);
  ahq = block_sum_ll(ahq);
  const unsigned qg = block_max_u(mq & 0xFFFFu), qh = block_max_u(mq >> 16);
  if (threadIdx.x == 0) {
    atomicAdd((unsigned long long*)&it->root_int_sum, (unsigned long long)ai);
    atomicAdd((unsigned long long*)&it->root_gq_sum, (unsigned long long)agq);
    atomicAdd((unsigned long long*)&it->root_hq_sum, (unsigned long long)ahq);
    atomicMax(&it->qmax_g, qg);
    atomicMax(&it->qmax_h, qh);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    it->grad_scale = gs;
    it->hess_scale = hs;
    it->start = (int)start;
  }
}

// A quantized tree on a full-data bag (not subset mode): the quantizer ran
// over every row into q_all (learner/quant.rs:92-129 with rows None); the
// root's positions take the bag rows' values, and the root sums are
// quant.rs:144-158 bag_root_sums: sums over the bag's rows of f32(gq * gs)
// and f32(hq * hs) with the scales rounded to f32, and the packed sum of
// positions 0 .. m - 1 (rows 0 .. m - 1, not the bag's). Every product is an
// integer multiple of 2^e (gs = M 2^e, M < 2^24), so the sums are exact
// int64 in those units; the host bounds m * bins so that the reference's
// f64 running sums are exact too.
__device__ __forceinline__ int f32_unit_exp(float s) { return s > 0.0f ? ilogbf(s) - 23 : 0; }

__global__ void __launch_bounds__(256) k_full_bag_root(long long m, const unsigned* __restrict__ bag,
                                                       const char2* __restrict__ q_all, IterDev* it,
                                                       char2* __restrict__ q0) {
  const float gs = (float)it->grad_scale, hs = (float)it->hess_scale;
  const int eg = f32_unit_exp(gs), eh = f32_unit_exp(hs);
  long long ag = 0, ah = 0, ai = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (long long)gridDim.x * blockDim.x) {
    const char2 v = q_all[bag[i]];
    q0[i] = v;
    const float pg = (float)v.x * gs, ph = (float)v.y * hs;
    ag += (long long)ldexp((double)pg, -eg);
    ah += (long long)ldexp((double)ph, -eh);
    const char2 u = q_all[i];
    ai += (long long)(((unsigned long long)(long long)u.x << 32) | (unsigned long long)(uint8_t)u.y);
  }
  ag = block_sum_ll(ag);
  ah = block_sum_ll(ah);
  ai = block_sum_ll(ai);
  if (threadIdx.x == 0) {
    atomicAdd((unsigned long long*)&it->bag_sg_fix, (unsigned long long)ag);
    atomicAdd((unsigned long long*)&it->bag_sh_fix, (unsigned long long)ah);
    atomicAdd((unsigned long long*)&it->bag_isum, (unsigned long long)ai);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    it->full_bag = 1;
    it->bag_eg = eg;
    it->bag_eh = eh;
  }
}

// ---------------------------------------------------------------- threshold search
// Shared scratch of one wave's search.
struct ScanLds {
  double val[3][kMaxBins + 1];   // g, h, count of each scan position (count as exact double)
  double pre[3][kMaxBins + 1];   // running sums in scan order
  int skipped[kMaxBins + 1];
};

struct LeafSums {
  double sg, sh;     // sh already includes 2 * kEpsilon
  int nd;
  double po;
  long long isum;    // quantized packed sum
  double gs, hs;
};

// One sequential scan of split.rs:676-891 on a wave: the prefix chains run
// in scan order on lanes 0-2 (g, h, count), every candidate is then scored
// in parallel and the first-in-scan-order maximum is kept.
__device__ __forceinline__ void scan_f64(const double* hg, const double* hh, const FeatDev& fd, const SplitParamsDev& p,
                         const LeafSums& ls, double min_gain_shift, bool reverse, bool skip_default, bool na,
                         ScanLds& S, SplitRec& out, bool& splittable) {
  const int lane = threadIdx.x & 63;
  const int off = fd.offset;
  const double cf = (double)ls.nd / ls.sh;
  int t_hi, t_lo, K;
  if (reverse) {
    t_hi = fd.num_bin - 1 - off - (na ? 1 : 0);
    t_lo = 1 - off;
    K = t_hi - t_lo + 1;
  } else {
    t_lo = (na && off == 1) ? -1 : 0;
    t_hi = fd.num_bin - 2 - off;
    K = t_hi - t_lo + 1;
  }
  if (K <= 0) return;
  // scan position k -> view index t -> bin t + off
  for (int k = lane; k < K; k += 64) {
    int t = reverse ? t_hi - k : t_lo + k;
    int skip = skip_default && (t + off) == fd.default_bin;
    S.skipped[k] = skip;
    if (t >= 0) {
      double hv = hh[t + off];
      S.val[0][k] = hg[t + off];
      S.val[1][k] = hv;
      S.val[2][k] = (double)round_int(hv * cf);
    } else {
      S.val[0][k] = S.val[1][k] = S.val[2][k] = 0.0;
    }
  }
  wave_sync();
  if (lane < 3) {
    double acc;
    const double* __restrict__ src = S.val[lane];
    double* __restrict__ dst = S.pre[lane];
    const int* __restrict__ skp = S.skipped;
    if (reverse) {
      acc = lane == 1 ? KEPS : 0.0;
    } else if (na && off == 1) {
      // split.rs:795-806: everything but the NaN bin starts on the left
      acc = lane == 0 ? ls.sg : lane == 1 ? ls.sh - KEPS : (double)ls.nd;
      const double* arr = lane == 0 ? hg : hh;
      const int n = fd.num_bin - off;
      int i = 0;
      for (; i + 8 <= n; i += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = lane == 2 ? (double)round_int(hh[i + u + off] * cf) : arr[i + u + off];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc -= v[u];
      }
      for (; i < n; ++i) acc -= lane == 2 ? (double)round_int(hh[i + off] * cf) : arr[i + off];
    } else {
      acc = lane == 1 ? KEPS : 0.0;
    }
    // the t >= 0 guard only matters at the forward scan's NaN position (t = -1, k = 0)
    const int k_first_real = (!reverse && t_lo < 0) ? 1 : 0;
    int k = 0;
    for (; k < k_first_real; ++k) dst[k] = acc;
    for (; k + 8 <= K; k += 8) {
      double v[8];
      int sk[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        v[u] = src[k + u];
        sk[u] = skp[k + u];
      }
      double o[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        acc = sk[u] ? acc : acc + v[u];
        o[u] = acc;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) dst[k + u] = o[u];
    }
    for (; k < K; ++k) {
      acc = skp[k] ? acc : acc + src[k];
      dst[k] = acc;
    }
  }
  wave_sync();
  // candidates: kb = first break in scan order; best = max gain before it
  int kb = INT32_MAX;
  double bg = -INFINITY;
  int bk = INT32_MAX;
  bool spl = false;
  for (int k = lane; k < K; k += 64) {
    if (S.skipped[k]) continue;
    double sa = S.pre[0][k], ha = S.pre[1][k];
    int ca = (int)S.pre[2][k];
    // reverse: a = right side; forward: a = left side
    if (ca < p.min_data_in_leaf || ha < p.min_sum_hessian) continue;
    int cb = ls.nd - ca;
    double hb = ls.sh - ha;
    if (cb < p.min_data_in_leaf || hb < p.min_sum_hessian) {
      kb = min(kb, k);
      continue;
    }
    double sb = ls.sg - sa;
    double gain = reverse ? split_gain(sb, hb, sa, ha, p, cb, ca, ls.po) : split_gain(sa, ha, sb, hb, p, ca, cb, ls.po);
    if (gain <= min_gain_shift) continue;
    // candidate at k (valid only if k < first break)
    S.skipped[k] = 2 + 0;  // marks a candidate; gains kept in val[0]
    S.val[0][k] = gain;
  }
  // first break
  kb = wave_min_i32(kb);
  for (int k = lane; k < K; k += 64) {
    if (S.skipped[k] == 2 && k < kb) {
      spl = true;
      double gain = S.val[0][k];
      if (gain > bg || (gain == bg && k < bk)) { bg = gain; bk = k; }
    }
  }
  wave_first_max(bg, bk);
  int any = __any(spl);
  wave_sync();
  if (!any) return;
  splittable = true;
  if (bk != INT32_MAX && bg > out.gain + min_gain_shift) {
    double sa = S.pre[0][bk], ha = S.pre[1][bk];
    int ca = (int)S.pre[2][bk];
    int t = reverse ? t_hi - bk : t_lo + bk;
    double slg, slh;
    int lc;
    if (reverse) { slg = ls.sg - sa; slh = ls.sh - ha; lc = ls.nd - ca; }
    else { slg = sa; slh = ha; lc = ca; }
    out.threshold = reverse ? t - 1 + off : t + off;
    out.left_output = leaf_output(slg, slh, p, lc, ls.po);
    out.left_count = lc;
    out.lsg = slg;
    out.lsh = slh - KEPS;
    out.right_output = leaf_output(ls.sg - slg, ls.sh - slh, p, ls.nd - lc, ls.po);
    out.right_count = ls.nd - lc;
    out.rsg = ls.sg - slg;
    out.rsh = ls.sh - slh - KEPS;
    out.gain = bg - min_gain_shift;
    out.default_left = reverse ? 1 : 0;
  }
}

// split.rs:402-441 on a wave.
__device__ __forceinline__ bool search_numerical(const double* hg, const double* hh, const FeatDev& fd, const SplitParamsDev& p,
                                 const LeafSums& ls, ScanLds& S, SplitRec& out) {
  double shift = leaf_gain(ls.sg, ls.sh, p, ls.nd, ls.po) + p.min_gain_to_split;
  bool spl = false;
  if (fd.num_bin > 2 && fd.missing_type != 0) {
    if (fd.missing_type == 1) {
      scan_f64(hg, hh, fd, p, ls, shift, true, true, false, S, out, spl);
      scan_f64(hg, hh, fd, p, ls, shift, false, true, false, S, out, spl);
    } else {
      scan_f64(hg, hh, fd, p, ls, shift, true, false, true, S, out, spl);
      scan_f64(hg, hh, fd, p, ls, shift, false, false, true, S, out, spl);
    }
  } else {
    scan_f64(hg, hh, fd, p, ls, shift, true, false, false, S, out, spl);
    if (fd.missing_type == 2) out.default_left = 0;
  }
  return spl;
}

// split.rs:446-670 (categorical, no monotone constraints) on lane 0; the
// ctr order is a stable rank sort computed by all lanes.
__device__ __forceinline__ bool search_categorical(const double* hg, const double* hh, const FeatDev& fd, const SplitParamsDev& p0,
                                   const LeafSums& ls, ScanLds& S, SplitRec& out) {
  const int lane = threadIdx.x & 63;
  const int off = fd.offset;
  const int bin_start = 1 - off, bin_end = fd.num_bin - off;
  const double cf = (double)ls.nd / ls.sh;
  SplitParamsDev pc = p0;
  const bool onehot = fd.num_bin <= p0.max_cat_to_onehot;
  double gain_shift = p0.path_smooth > KEPS ? 0.0 : leaf_gain(ls.sg, ls.sh, p0, ls.nd, 0.0);
  if (p0.path_smooth > KEPS) {
    double s = p0.lambda_l1 > 0.0 ? thr_l1(ls.sg, p0.lambda_l1) : ls.sg;
    gain_shift = -(2.0 * s * ls.po + (ls.sh + p0.lambda_l2) * ls.po * ls.po);
  }
  double min_gain_shift = gain_shift + p0.min_gain_to_split;
  out.default_left = 0;
  int used_bin = 0;
  int* sorted = (int*)S.pre[2];   // reuse scratch as ints
  if (!onehot) {
    // bins kept for the sort: round_int(h * cf) >= cat_smooth, in bin order
    for (int i = lane; i < bin_end - bin_start; i += 64) {
      int t = bin_start + i;
      S.skipped[i] = (double)round_int(hh[t + off] * cf) >= p0.cat_smooth;
    }
    wave_sync();
    if (lane == 0) {
      int c = 0;
      for (int i = 0; i < bin_end - bin_start; ++i)
        if (S.skipped[i]) sorted[c++] = bin_start + i;
      S.skipped[kMaxBins] = c;
    }
    wave_sync();
    used_bin = S.skipped[kMaxBins];
    // stable sort by ctr = g / (h + cat_smooth): rank = #less + #equal-before
    for (int i = lane; i < used_bin; i += 64) {
      int ti = sorted[i];
      double ci = hg[ti + off] / (hh[ti + off] + p0.cat_smooth);
      int rank = 0;
      for (int k = 0; k < used_bin; ++k) {
        int tk = sorted[k];
        double ck = hg[tk + off] / (hh[tk + off] + p0.cat_smooth);
        if (ck < ci || (!(ci < ck) && !(ck < ci) && k < i)) rank++;
      }
      S.val[2][rank] = (double)ti;
    }
    wave_sync();
  }
  bool is_splittable = false;
  if (lane == 0) {
    double best_gain = -INFINITY, best_slg = 0, best_slh = 0;
    int best_lc = 0, best_threshold = -1, best_dir = 1;
    if (onehot) {
      for (int t = bin_start; t < bin_end; ++t) {
        double grad = hg[t + off], hess = hh[t + off];
        int cnt = round_int(hess * cf);
        if (cnt < p0.min_data_in_leaf || hess < p0.min_sum_hessian) continue;
        int other = ls.nd - cnt;
        if (other < p0.min_data_in_leaf) continue;
        double soh = ls.sh - hess - KEPS;
        if (soh < p0.min_sum_hessian) continue;
        double sog = ls.sg - grad;
        double cg = split_gain(sog, soh, grad, hess + KEPS, pc, other, cnt, ls.po);
        if (cg <= min_gain_shift) continue;
        is_splittable = true;
        if (cg > best_gain) { best_threshold = t; best_slg = grad; best_slh = hess + KEPS; best_lc = cnt; best_gain = cg; }
      }
    } else {
      pc.lambda_l2 += p0.cat_l2;
      int max_num_cat = min(p0.max_cat_threshold, (used_bin + 1) / 2);
      for (int d = 0; d < 2; ++d) {
        int dir = d == 0 ? 1 : -1;
        int pos = d == 0 ? 0 : used_bin - 1;
        int cur_group = 0;
        double slg = 0.0, slh = KEPS;
        int lc = 0;
        for (int i = 0; i < used_bin && i < max_num_cat; ++i) {
          int t = (int)S.val[2][pos];
          pos += dir;
          double grad = hg[t + off], hess = hh[t + off];
          int cnt = round_int(hess * cf);
          slg += grad;
          slh += hess;
          lc += cnt;
          cur_group += cnt;
          if (lc < p0.min_data_in_leaf || slh < p0.min_sum_hessian) continue;
          int rc = ls.nd - lc;
          if (rc < p0.min_data_in_leaf || rc < p0.min_data_per_group) break;
          double srh = ls.sh - slh;
          if (srh < p0.min_sum_hessian) break;
          if (cur_group < p0.min_data_per_group) continue;
          cur_group = 0;
          double srg = ls.sg - slg;
          double cg = split_gain(slg, slh, srg, srh, pc, lc, rc, ls.po);
          if (cg <= min_gain_shift) continue;
          is_splittable = true;
          if (cg > best_gain) {
            best_lc = lc; best_slg = slg; best_slh = slh; best_threshold = i; best_gain = cg; best_dir = dir;
          }
        }
      }
    }
    if (is_splittable) {
      out.left_output = leaf_output(best_slg, best_slh, pc, best_lc, ls.po);
      out.left_count = best_lc;
      out.lsg = best_slg;
      out.lsh = best_slh - KEPS;
      out.right_output = leaf_output(ls.sg - best_slg, ls.sh - best_slh, pc, ls.nd - best_lc, ls.po);
      out.right_count = ls.nd - best_lc;
      out.rsg = ls.sg - best_slg;
      out.rsh = ls.sh - best_slh - KEPS;
      out.gain = best_gain - min_gain_shift;
      for (int w = 0; w < kCatWords; ++w) out.cat_bits[w] = 0;
      int nc = 0;
      if (onehot) {
        int b = best_threshold + off;
        out.cat_bits[b / 32] |= 1u << (b % 32);
        nc = 1;
      } else {
        for (int i = 0; i <= best_threshold; ++i) {
          int t = best_dir == 1 ? (int)S.val[2][i] : (int)S.val[2][used_bin - 1 - i];
          int b = t + off;
          out.cat_bits[b / 32] |= 1u << (b % 32);
          nc++;
        }
      }
      out.num_cat = nc;
    }
    S.skipped[kMaxBins] = is_splittable;
  }
  wave_sync();
  return S.skipped[kMaxBins] != 0;
}

// split.rs:1194-1400 (categorical_int, no monotone constraints) on lane 0:
// the categorical search on packed sums. Counts come from integer hessians
// (round_int(int_hess * nd / int_hess(total))); the chosen set's counts are
// recomputed from its packed sum; no kEpsilon. The ctr order is a stable
// rank sort computed by all lanes.
__device__ __forceinline__ bool search_categorical_int(const long long* hv, const FeatDev& fd, const SplitParamsDev& p0,
                                                       const LeafSums& ls, ScanLds& S, SplitRec& out) {
  const int lane = threadIdx.x & 63;
  const int off = fd.offset;
  const int bin_start = 1 - off, bin_end = fd.num_bin - off;
  const unsigned long long total = (unsigned long long)ls.isum;
  const double cf = (double)ls.nd / (double)(unsigned)total;
  const double sg = (double)(int)(ls.isum >> 32) * ls.gs, sh = (double)(unsigned)total * ls.hs;
  SplitParamsDev pc = p0;
  const bool onehot = fd.num_bin <= p0.max_cat_to_onehot;
  double gain_shift = p0.path_smooth > KEPS ? 0.0 : leaf_gain(sg, sh, p0, ls.nd, 0.0);
  if (p0.path_smooth > KEPS) {
    double s = p0.lambda_l1 > 0.0 ? thr_l1(sg, p0.lambda_l1) : sg;
    gain_shift = -(2.0 * s * ls.po + (sh + p0.lambda_l2) * ls.po * ls.po);
  }
  const double min_gain_shift = gain_shift + p0.min_gain_to_split;
  out.default_left = 0;
  int used_bin = 0;
  int* sorted = (int*)S.pre[2];   // scratch reused as ints
  if (!onehot) {
    for (int i = lane; i < bin_end - bin_start; i += 64) {
      const unsigned hq = (unsigned)hv[bin_start + i + off];
      S.skipped[i] = (double)round_int((double)hq * cf) >= p0.cat_smooth;
    }
    wave_sync();
    if (lane == 0) {
      int c = 0;
      for (int i = 0; i < bin_end - bin_start; ++i)
        if (S.skipped[i]) sorted[c++] = bin_start + i;
      S.skipped[kMaxBins] = c;
    }
    wave_sync();
    used_bin = S.skipped[kMaxBins];
    // stable sort by ctr = g / (h + cat_smooth): rank = #less + #equal-before
    for (int i = lane; i < used_bin; i += 64) {
      const int ti = sorted[i];
      const long long vi = hv[ti + off];
      const double ci = ((double)(int)(vi >> 32) * ls.gs) / ((double)(unsigned)vi * ls.hs + p0.cat_smooth);
      int rank = 0;
      for (int k = 0; k < used_bin; ++k) {
        const long long vk = hv[sorted[k] + off];
        const double ck = ((double)(int)(vk >> 32) * ls.gs) / ((double)(unsigned)vk * ls.hs + p0.cat_smooth);
        if (ck < ci || (!(ci < ck) && !(ck < ci) && k < i)) rank++;
      }
      S.val[2][rank] = (double)ti;
    }
    wave_sync();
  }
  bool is_splittable = false;
  if (lane == 0) {
    double best_gain = -INFINITY;
    unsigned long long best_left = 0;
    int best_threshold = -1, best_dir = 1;
    if (onehot) {
      for (int t = bin_start; t < bin_end; ++t) {
        const unsigned long long gh = (unsigned long long)hv[t + off];
        const int cnt = round_int((double)(unsigned)gh * cf);
        const double hess = (double)(unsigned)gh * ls.hs;
        if (cnt < p0.min_data_in_leaf || hess < p0.min_sum_hessian) continue;
        const int other = ls.nd - cnt;
        if (other < p0.min_data_in_leaf) continue;
        const unsigned long long so = total - gh;
        const double soh = (double)(unsigned)so * ls.hs;
        if (soh < p0.min_sum_hessian) continue;
        const double grad = (double)(int)(gh >> 32) * ls.gs;
        const double sog = (double)(int)(so >> 32) * ls.gs;
        const double cg = split_gain(sog, soh, grad, hess, pc, other, cnt, ls.po);
        if (cg <= min_gain_shift) continue;
        is_splittable = true;
        if (cg > best_gain) { best_threshold = t; best_left = gh; best_gain = cg; }
      }
    } else {
      pc.lambda_l2 += p0.cat_l2;
      const int max_num_cat = min(p0.max_cat_threshold, (used_bin + 1) / 2);
      for (int d = 0; d < 2; ++d) {
        const int dir = d == 0 ? 1 : -1;
        int pos = d == 0 ? 0 : used_bin - 1;
        int cur_group = 0, lc = 0;
        unsigned long long sl = 0;
        for (int i = 0; i < used_bin && i < max_num_cat; ++i) {
          const int t = (int)S.val[2][pos];
          pos += dir;
          const unsigned long long gh = (unsigned long long)hv[t + off];
          const int cnt = round_int((double)(unsigned)gh * cf);
          sl += gh;
          lc += cnt;
          cur_group += cnt;
          const double slh = (double)(unsigned)sl * ls.hs;
          if (lc < p0.min_data_in_leaf || slh < p0.min_sum_hessian) continue;
          const int rc = ls.nd - lc;
          if (rc < p0.min_data_in_leaf || rc < p0.min_data_per_group) break;
          const unsigned long long sr = total - sl;
          const double srh = (double)(unsigned)sr * ls.hs;
          if (srh < p0.min_sum_hessian) break;
          if (cur_group < p0.min_data_per_group) continue;
          cur_group = 0;
          const double slg = (double)(int)(sl >> 32) * ls.gs;
          const double srg = (double)(int)(sr >> 32) * ls.gs;
          const double cg = split_gain(slg, slh, srg, srh, pc, lc, rc, ls.po);
          if (cg <= min_gain_shift) continue;
          is_splittable = true;
          if (cg > best_gain) { best_left = sl; best_threshold = i; best_gain = cg; best_dir = dir; }
        }
      }
    }
    if (is_splittable) {
      const unsigned long long left = best_left, right = total - left;
      const double lg = (double)(int)(left >> 32) * ls.gs, lh = (double)(unsigned)left * ls.hs;
      const double rg = (double)(int)(right >> 32) * ls.gs, rh = (double)(unsigned)right * ls.hs;
      const int lc = round_int((double)(unsigned)left * cf), rc = round_int((double)(unsigned)right * cf);
      out.left_output = leaf_output(lg, lh, pc, lc, ls.po);
      out.left_count = lc;
      out.lsg = lg;
      out.lsh = lh;
      out.lgh = (long long)left;
      out.right_output = leaf_output(rg, rh, pc, rc, ls.po);
      out.right_count = rc;
      out.rsg = rg;
      out.rsh = rh;
      out.rgh = (long long)right;
      out.gain = best_gain - min_gain_shift;
      for (int w = 0; w < kCatWords; ++w) out.cat_bits[w] = 0;
      int nc = 0;
      if (onehot) {
        const int b = best_threshold + off;
        out.cat_bits[b / 32] |= 1u << (b % 32);
        nc = 1;
      } else {
        for (int i = 0; i <= best_threshold; ++i) {
          const int t = best_dir == 1 ? (int)S.val[2][i] : (int)S.val[2][used_bin - 1 - i];
          const int b = t + off;
          out.cat_bits[b / 32] |= 1u << (b % 32);
          nc++;
        }
      }
      out.num_cat = nc;
    }
    S.skipped[kMaxBins] = is_splittable;
  }
  wave_sync();
  return S.skipped[kMaxBins] != 0;
}

// split.rs:993-1190 on a wave: packed int64 prefix sums are exact, so every
// scan position is scored in parallel from a wave-parallel prefix.
__device__ __forceinline__ void scan_int(const long long* hv, const FeatDev& fd, const SplitParamsDev& p, const LeafSums& ls,
                         double min_gain_shift, bool reverse, bool skip_default, bool na, ScanLds& S, SplitRec& out,
                         bool& splittable) {
  const int lane = threadIdx.x & 63;
  const int off = fd.offset;
  long long* pre = (long long*)S.pre[0];
  int t_hi, t_lo, K;
  if (reverse) { t_hi = fd.num_bin - 1 - off - (na ? 1 : 0); t_lo = 1 - off; }
  else { t_lo = (na && off == 1) ? -1 : 0; t_hi = fd.num_bin - 2 - off; }
  K = t_hi - t_lo + 1;
  if (K <= 0) return;
  const double cf = (double)ls.nd / (double)(unsigned)ls.isum;
  // packed sums wrap exactly, so the prefix is a wave scan: lane l owns
  // positions 4l..4l+3
  unsigned long long base = 0;
  if (!reverse && na && off == 1) {
    unsigned long long part = 0;
    for (int i = lane; i < fd.num_bin - off; i += 64) part += (unsigned long long)hv[i + off];
    part = wave_incl_scan_u64(part);
    const unsigned long long total = ((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)(part >> 32), 63) << 32) |
                                     (unsigned)__builtin_amdgcn_readlane((int)(unsigned)part, 63);
    base = (unsigned long long)ls.isum - total;
  }
  unsigned long long v4[4], run = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int k = lane * 4 + j;
    unsigned long long v = 0;
    int skip = 1;
    if (k < K) {
      const int t = reverse ? t_hi - k : t_lo + k;
      skip = skip_default && (t + off) == fd.default_bin;
      if (!skip && t >= 0) v = (unsigned long long)hv[t + off];
      S.skipped[k] = skip;
    }
    run += v;
    v4[j] = run;
  }
  const unsigned long long incl = wave_incl_scan_u64(run);
  const unsigned long long excl = incl - run + base;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int k = lane * 4 + j;
    if (k < K) pre[k] = (long long)(excl + v4[j]);
  }
  wave_sync();
  int kb = INT32_MAX;
  for (int k = lane; k < K; k += 64) {
    if (S.skipped[k]) continue;
    long long a = pre[k];
    int ca = round_int((double)(unsigned)a * cf);
    double ha = (double)(unsigned)a * ls.hs;
    if (ca < p.min_data_in_leaf || ha < p.min_sum_hessian) continue;
    int cb = ls.nd - ca;
    if (cb < p.min_data_in_leaf) { kb = min(kb, k); continue; }
    long long b = (long long)((unsigned long long)ls.isum - (unsigned long long)a);
    double hb = (double)(unsigned)b * ls.hs;
    if (hb < p.min_sum_hessian) { kb = min(kb, k); continue; }
    double ga = (double)(int)(a >> 32) * ls.gs, gb = (double)(int)(b >> 32) * ls.gs;
    double gain = reverse ? split_gain(gb, hb + KEPS, ga, ha + KEPS, p, cb, ca, ls.po)
                          : split_gain(ga, ha + KEPS, gb, hb + KEPS, p, ca, cb, ls.po);
    if (gain <= min_gain_shift) continue;
    S.skipped[k] = 2;
    S.val[0][k] = gain;
  }
  kb = wave_min_i32(kb);
  double bg = -INFINITY;
  int bk = INT32_MAX;
  bool spl = false;
  for (int k = lane; k < K; k += 64) {
    if (S.skipped[k] == 2 && k < kb) {
      spl = true;
      double gain = S.val[0][k];
      if (gain > bg || (gain == bg && k < bk)) { bg = gain; bk = k; }
    }
  }
  wave_first_max(bg, bk);
  int any = __any(spl);
  wave_sync();
  if (!any) return;
  splittable = true;
  if (bk != INT32_MAX && bg > out.gain + min_gain_shift) {
    long long a = pre[bk];
    int t = reverse ? t_hi - bk : t_lo + bk;
    long long left = reverse ? (long long)((unsigned long long)ls.isum - (unsigned long long)a) : a;
    long long right = (long long)((unsigned long long)ls.isum - (unsigned long long)left);
    double lg = (double)(int)(left >> 32) * ls.gs, lh = (double)(unsigned)left * ls.hs;
    double rg = (double)(int)(right >> 32) * ls.gs, rh = (double)(unsigned)right * ls.hs;
    int lc = round_int((double)(unsigned)left * cf);
    int rc = round_int((double)(unsigned)right * cf);
    out.threshold = reverse ? t - 1 + off : t + off;
    out.left_output = leaf_output(lg, lh, p, lc, ls.po);
    out.left_count = lc;
    out.lsg = lg;
    out.lsh = lh;
    out.lgh = left;
    out.right_output = leaf_output(rg, rh, p, rc, ls.po);
    out.right_count = rc;
    out.rsg = rg;
    out.rsh = rh;
    out.rgh = right;
    out.gain = bg - min_gain_shift;
    out.default_left = reverse ? 1 : 0;
  }
}

__device__ __forceinline__ bool search_int(const long long* hv, const FeatDev& fd, const SplitParamsDev& p, const LeafSums& ls,
                           ScanLds& S, SplitRec& out) {
  double shift = leaf_gain((double)(int)(ls.isum >> 32) * ls.gs, (double)(unsigned)ls.isum * ls.hs, p, ls.nd, ls.po) +
                 p.min_gain_to_split;
  bool spl = false;
  if (fd.num_bin > 2 && fd.missing_type != 0) {
    if (fd.missing_type == 1) {
      scan_int(hv, fd, p, ls, shift, true, true, false, S, out, spl);
      scan_int(hv, fd, p, ls, shift, false, true, false, S, out, spl);
    } else {
      scan_int(hv, fd, p, ls, shift, true, false, true, S, out, spl);
      scan_int(hv, fd, p, ls, shift, false, false, true, S, out, spl);
    }
  } else {
    scan_int(hv, fd, p, ls, shift, true, false, false, S, out, spl);
    if (fd.missing_type == 2) out.default_left = 0;
  }
  return spl;
}

// ---------------------------------------------------------------- partition
// learner/mod.rs:792-810 goes_left / goes_left_categorical on a bin.
__device__ __forceinline__ bool goes_left(int b, const FeatDev& fd, const SplitRec& s) {
  if (fd.bin_type == 1) {
    int q = b;
    if (b == fd.mfb) {
      if (fd.mfb == 0) return false;
      q = fd.mfb;
    }
    return (s.cat_bits[q >> 5] >> (q & 31)) & 1u;
  }
  const bool miss = (fd.missing_type == 1 && b == fd.default_bin) || (fd.missing_type == 2 && b == fd.num_bin - 1);
  return miss ? s.default_left != 0 : b <= s.threshold;
}

// ---------------------------------------------------------------- prediction
// Flattened trees: the internal nodes of all trees concatenated, 16 bytes each.
//   x = feature (16 bits) | flags << 16 | thr_bin << 24
//       flags: 1 default_left, 2 missing type Zero, 4 missing type NaN, 8 categorical
//   y = left (int16) | right (int16) << 16; a child >= 0 is a node, < 0 is ~leaf
//   z, w = raw numerical: the f64 threshold's bits (low, high);
//          categorical: offset and length of the node's bitset words;
//          binned numerical: z = the feature's missing bin (default bin for
//          Zero, last bin for NaN).
// The feature is the raw column for raw rows and the inner feature for bins.
struct PredArgs {
  const uint4* nodes;
  const int* tree_node_off;   // [T]
  const int* tree_leaf_off;   // [T]
  const int* tree_num_leaves; // [T]
  const double* leaf_val;
  const unsigned* cat_words;
  int T, K;
  long long rows;
  int ncol;                   // columns of a row: raw columns, or F for bins
  int stride;                 // staged row stride in LDS (elements)
  const float* X;             // raw rows (row-major f32)
  const uint8_t* bins;        // binned rows (row-major u8)
  double* out;                // [rows][K]
  int accum;                  // 1: out[row][k] += the trees' sum (one tree: its leaf value, exactly)
};

constexpr int kPredRows = 64;   // rows per block (one wave, one row per lane)
#ifndef LGBMHIP_PRED_TREES
#define LGBMHIP_PRED_TREES 4
#endif
constexpr int kPredTrees = LGBMHIP_PRED_TREES;   // trees of one class walked together per lane

// One decision of a node (predict.rs:95-109 numerical step, tree.rs:399-418
// categorical; tree.rs:605-632 on bins).
template <bool BINNED>
__device__ __forceinline__ int pred_step(const uint4 n, const float* xr, const uint8_t* br, int ncol,
                                         const unsigned* cat_words) {
  const int f = n.x & 0xFFFF;
  const unsigned flags = (n.x >> 16) & 0xFF;
  const int left = (int)(short)(n.y & 0xFFFF), right = (int)(short)(n.y >> 16);
  bool go_left;
  if (BINNED) {
    const int b = br[f];
    if (flags & 8) {
      go_left = (b >> 5) < (int)n.w && ((cat_words[n.z + (b >> 5)] >> (b & 31)) & 1u);
    } else {
      const bool is_default = (flags & 6) && b == (int)n.z;
      go_left = is_default ? (flags & 1) != 0 : b <= (int)(n.x >> 24);
    }
  } else {
    double v = f < ncol ? (double)xr[f] : 0.0;
    if (flags & 8) {
      go_left = false;
      if (!isnan(v)) {
        const int c = (int)v;  // saturating, as the reference's `as i32`
        go_left = c >= 0 && (c >> 5) < (int)n.w && ((cat_words[n.z + (c >> 5)] >> (c & 31)) & 1u);
      }
    } else {
      const bool nan = isnan(v);
      // the reference's predictor drops |v| <= kZeroThreshold (predict.rs:262-267)
      if ((nan && !(flags & 4)) || (v >= -KZERO && v <= KZERO)) v = 0.0;
      const bool zero = v == 0.0;
      const bool is_default = ((flags & 2) && zero) || ((flags & 4) && nan);
      const double thr = __hiloint2double((int)n.w, (int)n.z);
      go_left = is_default ? (flags & 1) != 0 : v <= thr;
    }
  }
  return go_left ? left : right;
}

// Gbdt::predict_matrix_early_stop with early stopping (predict.rs:231-339,
// the margin test :155-171): one row per lane, the rows staged in LDS, the
// iterations in order (each class's tree added to the row's raw score of
// that class, as the reference adds them); every `freq` iterations a row
// whose margin exceeds the threshold stops taking trees.
template <bool BINNED>
__global__ void __launch_bounds__(kPredRows) k_predict_es(PredArgs a, int freq, double margin) {
  extern __shared__ __align__(16) unsigned char pred_lds[];
  __shared__ double raw[16][kPredRows];
  const long long r0 = (long long)blockIdx.x * kPredRows;
  const int nrow = (int)min((long long)kPredRows, a.rows - r0);
  const int lane = threadIdx.x;
  {
    const long long n = (long long)nrow * a.ncol;
    for (long long e = lane; e < n; e += kPredRows) {
      const int r = (int)(e / a.ncol), f = (int)(e % a.ncol);
      if (BINNED) pred_lds[r * a.stride + f] = a.bins[r0 * a.ncol + e];
      else ((float*)pred_lds)[r * a.stride + f] = a.X[r0 * a.ncol + e];
    }
  }
  for (int k = 0; k < a.K; ++k) raw[k][lane] = 0.0;
  __syncthreads();
  const int rr = min(lane, nrow - 1);
  const uint8_t* br = BINNED ? pred_lds + rr * a.stride : nullptr;
  const float* xr = BINNED ? nullptr : (const float*)pred_lds + rr * a.stride;
  bool stopped = lane >= nrow;
  const int iters = a.T / a.K;
  for (int it = 0; it < iters; ++it) {
    if (!stopped) {
      for (int k = 0; k < a.K; ++k) {
        const int t = it * a.K + k;
        int node = a.tree_num_leaves[t] > 1 ? 0 : ~0;
        const int base = a.tree_node_off[t];
        while (node >= 0) node = pred_step<BINNED>(a.nodes[base + node], xr, br, a.ncol, a.cat_words);
        raw[k][lane] += a.leaf_val[a.tree_leaf_off[t] + ~node];
      }
      if ((it + 1) % freq == 0) {
        double m;
        if (a.K > 1) {
          double x = -INFINITY, y = -INFINITY;
          for (int k = 0; k < a.K; ++k) {
            const double v = raw[k][lane];
            if (v > x) {
              y = x;
              x = v;
            } else if (v > y) {
              y = v;
            }
          }
          m = x - y;
        } else {
          m = 2.0 * fabs(raw[0][lane]);
        }
        if (m > margin) stopped = true;
      }
    }
  }
  if (lane < nrow)
    for (int k = 0; k < a.K; ++k) a.out[(r0 + lane) * a.K + k] = raw[k][lane];
}

// PredictKind::LeafIndex (predict.rs:272-279): each row's leaf in every tree
// of the window, out[row][tree] as f64; one row per lane, rows in LDS.
template <bool BINNED>
__global__ void __launch_bounds__(kPredRows) k_predict_leaf(PredArgs a) {
  extern __shared__ __align__(16) unsigned char pred_lds[];
  const long long r0 = (long long)blockIdx.x * kPredRows;
  const int nrow = (int)min((long long)kPredRows, a.rows - r0);
  const int lane = threadIdx.x;
  const long long n = (long long)nrow * a.ncol;
  for (long long e = lane; e < n; e += kPredRows) {
    const int r = (int)(e / a.ncol), f = (int)(e % a.ncol);
    if (BINNED) pred_lds[r * a.stride + f] = a.bins[r0 * a.ncol + e];
    else ((float*)pred_lds)[r * a.stride + f] = a.X[r0 * a.ncol + e];
  }
  __syncthreads();
  if (lane >= nrow) return;
  const uint8_t* br = BINNED ? pred_lds + lane * a.stride : nullptr;
  const float* xr = BINNED ? nullptr : (const float*)pred_lds + lane * a.stride;
  for (int t = 0; t < a.T; ++t) {
    int node = a.tree_num_leaves[t] > 1 ? 0 : ~0;
    const int base = a.tree_node_off[t];
    while (node >= 0) node = pred_step<BINNED>(a.nodes[base + node], xr, br, a.ncol, a.cat_words);
    a.out[(r0 + lane) * (long long)a.T + t] = (double)~node;
  }
}

// Gbdt::predict_matrix_early_stop without early stopping (predict.rs:231-339):
// one row per lane, the block's rows staged in LDS (STAGE) at an odd word
// stride so lanes reading one feature hit distinct banks. For each class,
// the class's trees are summed in model order; kPredTrees of them are walked
// together so their node loads overlap.
template <bool BINNED, bool STAGE, int W>
__global__ void __launch_bounds__(kPredRows * W) k_predict(PredArgs a) {
  // W waves share one 64-row tile: wave w walks trees w*G .. w*G + G - 1 of
  // every chunk of W*G trees of a class and leaves their values in LDS, then
  // wave 0 adds the chunk's values to each row's sum in model order.
  constexpr int G = kPredTrees;
  constexpr int TC = W * G;
  extern __shared__ __align__(16) unsigned char pred_lds[];
  __shared__ double vals[W > 1 ? TC : 1][kPredRows];
  const long long r0 = (long long)blockIdx.x * kPredRows;
  const int nrow = (int)min((long long)kPredRows, a.rows - r0);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int NT = kPredRows * W;
  const float* xr = nullptr;
  const uint8_t* br = nullptr;
  if (STAGE) {
    // copy the block's contiguous rows, element e = (r, f) -> r * stride + f
    const long long n = (long long)nrow * a.ncol;
    int r = 0, f = tid;
    while (f >= a.ncol) { f -= a.ncol; ++r; }
    const int step_r = NT / a.ncol, step_f = NT % a.ncol;
    for (long long e = tid; e < n; e += NT) {
      if (BINNED) pred_lds[r * a.stride + f] = a.bins[r0 * a.ncol + e];
      else ((float*)pred_lds)[r * a.stride + f] = a.X[r0 * a.ncol + e];
      r += step_r;
      f += step_f;
      if (f >= a.ncol) { f -= a.ncol; ++r; }
    }
    __syncthreads();
    const int rr = min(lane, nrow - 1);
    if (BINNED) br = pred_lds + rr * a.stride;
    else xr = (const float*)pred_lds + rr * a.stride;
  } else {
    const long long rr = r0 + min(lane, nrow - 1);
    if (BINNED) br = a.bins + rr * a.ncol;
    else xr = a.X + rr * a.ncol;
  }
  for (int k = 0; k < a.K; ++k) {
    // -0.0 + v == v for every v, so an accumulated single tree adds its leaf value unchanged
    double acc = a.accum ? -0.0 : 0.0;
    const int nk = a.T > k ? (a.T - k + a.K - 1) / a.K : 0;   // trees of class k
    for (int c0 = 0; c0 < nk; c0 += TC) {
      int node[G], base[G];
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int ci = c0 + wave * G + g;
        const int t = k + ci * a.K;
        const bool walk = ci < nk && a.tree_num_leaves[t] > 1;
        node[g] = walk ? 0 : ~0;
        base[g] = ci < nk ? a.tree_node_off[t] : 0;
      }
      bool more = true;
      while (more) {
        more = false;
#pragma unroll
        for (int g = 0; g < G; ++g) {
          if (node[g] >= 0) {
            const uint4 n = a.nodes[base[g] + node[g]];
            node[g] = pred_step<BINNED>(n, xr, br, a.ncol, a.cat_words);
            more |= node[g] >= 0;
          }
        }
      }
      if (W == 1) {
#pragma unroll
        for (int g = 0; g < G; ++g) {
          const int ci = c0 + g;
          if (ci < nk) acc += a.leaf_val[a.tree_leaf_off[k + ci * a.K] + ~node[g]];
        }
      } else {
#pragma unroll
        for (int g = 0; g < G; ++g) {
          const int ci = c0 + wave * G + g;
          if (ci < nk) vals[wave * G + g][lane] = a.leaf_val[a.tree_leaf_off[k + ci * a.K] + ~node[g]];
        }
        __syncthreads();
        if (wave == 0) {
          const int m = min(TC, nk - c0);
          for (int i = 0; i < m; ++i) acc += vals[i][lane];
        }
        __syncthreads();
      }
    }
    if (wave == 0 && lane < nrow) {
      double* o = a.out + (r0 + lane) * a.K + k;
      *o = a.accum ? *o + acc : acc;
    }
  }
}

// Stage entry points: stable partition of explicit rows (partition.rs:107-174)
// in tiles of kSpartTile rows, 256 threads of 4 consecutive rows each.
//   k_spart_count   one block per tile: the tile's left rows -> cnt[tile];
//                   a row index >= n sets cnt[ntiles + 1] (nothing is read for it)
//   k_spart_scan    one block: cnt[t] -> first left slot of tile t (exclusive
//                   prefix), cnt[ntiles] = all left rows
//   k_spart_scatter one block per tile: left rows in order from the tile's
//                   slot, right rows in order after all left rows
// Work buffer: ntiles + 2 words; cnt[ntiles + 1] must be zero before the count.
constexpr int kSpartTile = 1024;
// lgbmhip_partition_device_noalloc's work buffer: 256 KiB per device, enough
// for the tile counts of (kNoAllocScratchWords - 2) tiles of rows
constexpr int kNoAllocScratchWords = 1 << 16;
constexpr long long kNoAllocMaxRows = (long long)(kNoAllocScratchWords - 2) * kSpartTile;

// Stage entry point: score[i] += leaf_values[leaf[i]] (partition.rs:178-191),
// one add per row, so the order of rows does not matter.
__global__ void k_add_leaf_outputs(const int* __restrict__ leaf, const double* __restrict__ lv,
                                   double* __restrict__ score, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    score[i] += lv[leaf[i]];
}

// words p[0..n) = v (a kernel, so a caller's graph capture records no memset node)
__global__ void k_fill_u32(unsigned* p, long long n, unsigned v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

__device__ __forceinline__ void spart_load(const unsigned* rows, long long m, unsigned n, const uint8_t* col,
                                           const FeatDev& fd, const SplitRec& s, long long first, unsigned r[4],
                                           bool valid[4], bool left[4], bool* bad) {
  unsigned b[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    valid[u] = first + u < m;
    r[u] = valid[u] ? rows[first + u] : 0u;
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const bool in = valid[u] && r[u] < n;
    if (valid[u] && !in) *bad = true;
    b[u] = in ? col[r[u]] : 0u;
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) left[u] = valid[u] && r[u] < n && goes_left((int)b[u], fd, s);
}

__global__ void __launch_bounds__(256) k_spart_count(const unsigned* rows, long long m, unsigned n,
                                                     const uint8_t* col, FeatDev fd, SplitRec s, unsigned* cnt,
                                                     int ntiles) {
  __shared__ unsigned tot;
  if (threadIdx.x == 0) tot = 0;
  __syncthreads();
  unsigned r[4];
  bool valid[4], left[4], bad = false;
  spart_load(rows, m, n, col, fd, s, (long long)blockIdx.x * kSpartTile + threadIdx.x * 4, r, valid, left, &bad);
  unsigned c = (unsigned)left[0] + (unsigned)left[1] + (unsigned)left[2] + (unsigned)left[3];
  for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(&tot, c);
  if (bad) atomicOr(&cnt[ntiles + 1], 1u);
  __syncthreads();
  if (threadIdx.x == 0) cnt[blockIdx.x] = tot;
}

// Exclusive scan of cnt[0..ntiles) in place by one 1024-thread block, in
// passes of 1024 entries with a running carry; cnt[ntiles] = the total.
__global__ void __launch_bounds__(1024) k_spart_scan(unsigned* cnt, int ntiles) {
  __shared__ unsigned wsum[16];
  __shared__ unsigned carry;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  if (t == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < ntiles; base += 1024) {
    const int i = base + t;
    const unsigned v = i < ntiles ? cnt[i] : 0u;
    unsigned x = v;   // inclusive wave scan
    for (int o = 1; o < 64; o <<= 1) {
      const unsigned y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    unsigned before = carry;
    for (int k = 0; k < w; ++k) before += wsum[k];
    if (i < ntiles) cnt[i] = before + x - v;
    __syncthreads();
    if (t == 0) {
      unsigned add = 0;
      for (int k = 0; k < 16; ++k) add += wsum[k];
      carry += add;
    }
    __syncthreads();
  }
  if (t == 0) cnt[ntiles] = carry;
}

__global__ void __launch_bounds__(256) k_spart_scatter(const unsigned* rows, long long m, unsigned n,
                                                       const uint8_t* col, FeatDev fd, SplitRec s,
                                                       const unsigned* cnt, int ntiles, unsigned* out) {
  __shared__ unsigned wsum[4];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const long long tile0 = (long long)blockIdx.x * kSpartTile;
  unsigned r[4];
  bool valid[4], left[4], bad = false;
  spart_load(rows, m, n, col, fd, s, tile0 + t * 4, r, valid, left, &bad);
  const unsigned c = (unsigned)left[0] + (unsigned)left[1] + (unsigned)left[2] + (unsigned)left[3];
  unsigned x = c;   // inclusive wave scan of the threads' left counts
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  unsigned lbefore = x - c;   // left rows of this tile before this thread's rows
  for (int k = 0; k < w; ++k) lbefore += wsum[k];
  const unsigned tile_left = cnt[blockIdx.x], total_left = cnt[ntiles];
  unsigned lslot = tile_left + lbefore;
  // right rows before this thread's: the tile's earlier rows minus its earlier left rows
  unsigned rslot = total_left + ((unsigned)tile0 - tile_left) + ((unsigned)t * 4u - lbefore);
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    if (!valid[u]) continue;
    if (left[u]) out[lslot++] = r[u];
    else out[rslot++] = r[u];
  }
}

// Stage entry point: one feature's threshold search on a given histogram.
__global__ void __launch_bounds__(64) k_search_one(const double* hist, const long long* ihist, FeatDev fd,
                                                   SplitParamsDev p, LeafSums ls, int quant, SplitRec* out,
                                                   int* splittable) {
  __shared__ double hg[kMaxBins], hh[kMaxBins];
  __shared__ long long ih[kMaxBins];
  __shared__ ScanLds S;
  __shared__ SplitRec r;
  const int lane = threadIdx.x;
  for (int b = lane; b < kMaxBins; b += 64) {
    if (quant) ih[b] = b < fd.num_bin ? ihist[b] : 0;
    else {
      hg[b] = b < fd.num_bin ? hist[2 * b] : 0.0;
      hh[b] = b < fd.num_bin ? hist[2 * b + 1] : 0.0;
    }
  }
  if (lane == 0) reset_rec(r);
  __syncthreads();
  if (quant) {
    if (fd.mfb > 0 && lane == 0) {
      unsigned long long s = (unsigned long long)ls.isum;
      for (int i = 0; i < fd.num_bin; ++i) if (i != fd.mfb) s -= (unsigned long long)ih[i];
      ih[fd.mfb] = (long long)s;
    }
  } else if (fd.mfb > 0 && lane < 2) {
    double v = lane == 0 ? ls.sg : ls.sh - 2.0 * KEPS;
    double* arr = lane == 0 ? hg : hh;
    for (int i = 0; i < fd.num_bin; ++i) if (i != fd.mfb) v -= arr[i];
    arr[fd.mfb] = v;
  }
  __syncthreads();
  bool spl;
  if (quant) spl = fd.bin_type == 1 ? search_categorical_int(ih, fd, p, ls, S, r) : search_int(ih, fd, p, ls, S, r);
  else if (fd.bin_type == 1) spl = search_categorical(hg, hh, fd, p, ls, S, r);
  else spl = search_numerical(hg, hh, fd, p, ls, S, r);
  if (lane == 0) {
    *out = r;
    *splittable = spl ? 1 : 0;
  }
}

// ---------------------------------------------------------------- bagging
// sample_strategy.rs:285-318: row i is in the bag when the next draw of the
// generator of its 1024-row block (Random(seed + block), random.rs:15-40,
// its state kept in lcg[block] from draw to draw) is below the fraction.
// One thread per block draws its rows in order.
constexpr int kBagBlock = 1024;

// LCG jump-ahead: the map x -> a x + c applied k times is x -> A x + C (mod 2^32).
__device__ __forceinline__ void lcg_jump(unsigned k, unsigned& A, unsigned& C) {
  unsigned ma = 214013u, mc = 2531011u;   // the map applied 2^i times
  A = 1u;
  C = 0u;
  while (k) {
    if (k & 1u) {
      A = ma * A;
      C = ma * C + mc;
    }
    mc = ma * mc + mc;
    ma = ma * ma;
    k >>= 1;
  }
}

// One wave per generator block: lane j draws rows 16 j .. 16 j + 15 of the
// block from the generator state jumped ahead by 16 j steps, so the draws are
// the sequential ones; lane 0 stores the state after the block's rows.
__global__ void __launch_bounds__(256) k_bag_draw(long long n, unsigned* lcg, double fraction, const float* label,
                                                  double pos_fraction, double neg_fraction, uint8_t* flags,
                                                  int* counts) {
  const int lane = threadIdx.x & 63;
  const long long b = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long long nb = (n + kBagBlock - 1) / kBagBlock;
  if (b >= nb) return;
  const unsigned x0 = lcg[b];
  const long long r0 = b * kBagBlock, rows = min((long long)kBagBlock, n - r0);
  constexpr int PER = kBagBlock / 64;
  unsigned A, C;
  lcg_jump((unsigned)(lane * PER), A, C);
  unsigned x = A * x0 + C;
  int cnt = 0;
  for (int t = 0; t < PER; ++t) {
    const long long i = (long long)lane * PER + t;
    if (i >= rows) break;
    x = x * 214013u + 2531011u;
    const float f = (float)(int)((x >> 16) & 0x7FFFu) / 32768.0f;
    // balanced bagging: the fraction of the row's class (label > 0 or not)
    const double frac = label ? (label[r0 + i] > 0.0f ? pos_fraction : neg_fraction) : fraction;
    const int in = (double)f < frac ? 1 : 0;
    flags[r0 + i] = (uint8_t)in;
    cnt += in;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
  if (lane == 0) {
    lcg_jump((unsigned)rows, A, C);
    lcg[b] = A * x0 + C;
    counts[b] = cnt;
  }
}

// bagging_by_query (sample_strategy.rs:239-276): query q is in the bag when
// the next draw of the generator of its 1024-query block is below the
// fraction, and then all of its rows are. One wave per generator block: lane
// j draws queries 16 j .. 16 j + 15 of the block from the state jumped ahead
// by 16 j steps and flags their rows (every row is some query's).
__global__ void __launch_bounds__(256) k_bag_draw_queries(long long nq, unsigned* lcg, double fraction,
                                                          const int* __restrict__ qb, uint8_t* __restrict__ flags) {
  const int lane = threadIdx.x & 63;
  const long long b = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long long nb = (nq + kBagBlock - 1) / kBagBlock;
  if (b >= nb) return;
  const unsigned x0 = lcg[b];
  const long long q0 = b * kBagBlock, qs = min((long long)kBagBlock, nq - q0);
  constexpr int PER = kBagBlock / 64;
  unsigned A, C;
  lcg_jump((unsigned)(lane * PER), A, C);
  unsigned x = A * x0 + C;
  for (int t = 0; t < PER; ++t) {
    const long long i = (long long)lane * PER + t;
    if (i >= qs) break;
    x = x * 214013u + 2531011u;
    const float f = (float)(int)((x >> 16) & 0x7FFFu) / 32768.0f;
    const uint8_t in = (double)f < fraction ? 1 : 0;
    for (int r = qb[q0 + i]; r < qb[q0 + i + 1]; ++r) flags[r] = in;
  }
  if (lane == 0) {
    lcg_jump((unsigned)qs, A, C);
    lcg[b] = A * x0 + C;
  }
}

// Exclusive offsets of the per-block counts (one 1024-thread block), and the bag size.
__global__ void __launch_bounds__(1024) k_bag_offsets(const int* counts, long long nb, int* offsets, long long* total) {
  __shared__ int sm[1024];
  __shared__ int carry;
  const int t = threadIdx.x;
  if (t == 0) carry = 0;
  __syncthreads();
  for (long long base = 0; base < nb; base += 1024) {
    const long long i = base + t;
    const int v = i < nb ? counts[i] : 0;
    sm[t] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      const int add = t >= o ? sm[t - o] : 0;
      __syncthreads();
      sm[t] += add;
      __syncthreads();
    }
    if (i < nb) offsets[i] = carry + sm[t] - v;
    __syncthreads();
    if (t == 1023) carry += sm[1023];
    __syncthreads();
  }
  if (t == 0) *total = carry;
}

// In-bag rows ascending: one 1024-thread block per generator block.
__global__ void __launch_bounds__(1024) k_bag_scatter(long long n, const uint8_t* flags, const int* offsets,
                                                      unsigned* rows) {
  __shared__ int sm[1024];
  const int t = threadIdx.x;
  const long long i = (long long)blockIdx.x * kBagBlock + t;
  const int v = i < n ? flags[i] : 0;
  sm[t] = v;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int add = t >= o ? sm[t - o] : 0;
    __syncthreads();
    sm[t] += add;
    __syncthreads();
  }
  if (v) rows[offsets[blockIdx.x] + sm[t] - 1] = (unsigned)i;
}

// ---------------------------------------------------------------- GOSS
// sample_strategy.rs:395-442 on the device. |g h| summed over the K classes
// in f32, class by class.
__global__ void __launch_bounds__(256) k_goss_abs(const float* g, const float* h, long long n, int K, float* out) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    float s = 0.0f;
    for (int k = 0; k < K; ++k) s += fabsf(g[(size_t)k * n + i] * h[(size_t)k * n + i]);
    out[i] = s;
  }
}

// Per GOSS block (one 1024-thread workgroup each): the top_k-th largest
// |g h| of rows start .. start + cnt (ArgMaxAtK, array_args.rs:65-79), by
// a most-significant-digit radix select over the f32 bits (non-negative
// floats order as their bits).
struct GossBlock {
  long long start, cnt;
  int top_k, other_k;
  float multiply;
  int pad;
};
__global__ void __launch_bounds__(1024) k_goss_threshold(const float* v, const GossBlock* blocks, float* thr) {
  __shared__ unsigned hist[256];
  __shared__ unsigned s_prefix, s_k;
  const GossBlock gb = blocks[blockIdx.x];
  const int t = threadIdx.x;
  if (t == 0) {
    s_prefix = 0u;
    s_k = (unsigned)gb.top_k;   // rank from the top still to find
  }
  for (int shift = 24; shift >= 0; shift -= 8) {
    if (t < 256) hist[t] = 0u;
    __syncthreads();
    const unsigned prefix = s_prefix;
    const unsigned mask_hi = shift == 24 ? 0u : (0xFFFFFFFFu << (shift + 8));
    for (long long i = t; i < gb.cnt; i += blockDim.x) {
      const unsigned x = __float_as_uint(v[gb.start + i]);
      if ((x & mask_hi) == prefix) atomicAdd(&hist[(x >> shift) & 0xFFu], 1u);
    }
    __syncthreads();
    if (t == 0) {
      unsigned k = s_k;
      int d = 255;
      for (; d > 0; --d) {
        if (hist[d] >= k) break;
        k -= hist[d];
      }
      s_prefix = prefix | ((unsigned)d << shift);
      s_k = k;
    }
    __syncthreads();
  }
  if (t == 0) thr[blockIdx.x] = __uint_as_float(s_prefix);
}

// One wave per GOSS block: rows in 1024-row chunks (the generator blocks;
// GOSS blocks are multiples of 1024 rows). Lanes mark the rows at or above
// the threshold and draw, for every other row, the next value of its chunk's
// generator (jumped ahead by the chunk's earlier non-top rows); lane 0 then
// runs the reference's sequential sampling (rest_need / rest_all, compared
// exactly in integers: draw k / 32768 < a / b iff k b < 32768 a); sampled
// rows' g, h of every class are scaled by the block's multiplier.
__global__ void __launch_bounds__(64) k_goss_chain(const float* v, const GossBlock* blocks, const float* thr,
                                                   unsigned* lcg, float* g, float* h, long long n, int K,
                                                   uint8_t* flags) {
  __shared__ __align__(16) unsigned short draw[kBagBlock];   // 15-bit draw, 0xFFFF a top row, 0xFFFE no row
  __shared__ __align__(16) uint8_t take[kBagBlock];
  const GossBlock gb = blocks[blockIdx.x];
  const float th = thr[blockIdx.x];
  const int lane = threadIdx.x;
  constexpr int PER = kBagBlock / 64;
  int big = 0, left = 0;   // lane 0's running counts over the block
  for (long long c0 = 0; c0 < gb.cnt; c0 += kBagBlock) {
    const long long base = gb.start + c0;
    const int rows = (int)min((long long)kBagBlock, gb.cnt - c0);
    const long long gen = base / kBagBlock;
    bool top[PER];
    int nb = 0;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int i = lane * PER + j;
      top[j] = i < rows && v[base + i] >= th;
      nb += (i < rows && !top[j]) ? 1 : 0;
    }
    int incl = nb;
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(incl, o);
      if (lane >= o) incl += y;
    }
    const int before = incl - nb, total = __shfl(incl, 63);
    const unsigned x0 = lcg[gen];
    unsigned A, C;
    lcg_jump((unsigned)before, A, C);
    unsigned x = A * x0 + C;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int i = lane * PER + j;
      if (i >= rows) {
        draw[i] = 0xFFFEu;   // past the block's rows
      } else if (top[j]) {
        draw[i] = 0xFFFFu;
      } else {
        x = x * 214013u + 2531011u;
        draw[i] = (unsigned short)((x >> 16) & 0x7FFFu);
      }
    }
    __syncthreads();
    if (lane == 0) {
      lcg_jump((unsigned)total, A, C);
      lcg[gen] = A * x0 + C;
      // 8 draws per LDS read; rest_all counts down with the rows
      long long rest_base = gb.cnt - c0 - (long long)gb.top_k;
      for (int i0 = 0; i0 < rows; i0 += 8) {
        const uint4 q = *(const uint4*)&draw[i0];
        const unsigned w4[4] = {q.x, q.y, q.z, q.w};
        unsigned long long tk = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const unsigned d = (w4[j >> 1] >> (16 * (j & 1))) & 0xFFFFu;
          unsigned long long in = 0;
          if (d == 0xFFFFu) {
            in = 1;
            ++left;
            ++big;
          } else if (d != 0xFFFEu) {
            const long long rest_need = (long long)gb.other_k - (left - big);
            const long long rest_all = rest_base - (i0 + j) + big;
            in = (long long)d * rest_all < rest_need * 32768LL ? 2 : 0;
            left += in ? 1 : 0;
          }
          tk |= in << (8 * j);
        }
        *(unsigned long long*)&take[i0] = tk;
      }
    }
    __syncthreads();
    for (int i = lane; i < rows; i += 64) {
      const uint8_t in = take[i];
      flags[base + i] = in ? 1 : 0;
      if (in == 2)
        for (int k = 0; k < K; ++k) {
          g[(size_t)k * n + base + i] *= gb.multiply;
          h[(size_t)k * n + base + i] *= gb.multiply;
        }
    }
    __syncthreads();
  }
}

// Rows per generator block that the flags put in the bag (for k_bag_offsets).
__global__ void __launch_bounds__(256) k_bag_counts(const uint8_t* flags, long long n, int* counts) {
  const int lane = threadIdx.x & 63;
  const long long b = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long long nb = (n + kBagBlock - 1) / kBagBlock;
  if (b >= nb) return;
  int cnt = 0;
  for (long long i = b * kBagBlock + lane; i < min(n, (b + 1) * kBagBlock); i += 64) cnt += flags[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
  if (lane == 0) counts[b] = cnt;
}

// Clears class k's max |g|, |h| before they are taken over a bag.
__global__ void k_clear_max(IterDev* it) {
  it->max_abs_g_bits = 0;
  it->max_abs_h_bits = 0;
  it->qmax_g = 0;
  it->qmax_h = 0;
}

// ---------------------------------------------------------------- LDS-atomic peak
// Conflict-free LDS atomics: lane l of each wave adds to its own word, so a
// wave's 64 adds spread evenly over the banks. The block's words are summed
// into sink so that no add is dead.
__global__ void __launch_bounds__(512) k_lds_atomic_peak(int iters, unsigned* sink) {
  __shared__ unsigned cells[512];
  cells[threadIdx.x] = 0u;
  __syncthreads();
  unsigned* cell = cells + threadIdx.x;
  const unsigned v = threadIdx.x | 1u;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int u = 0; u < 16; ++u) __hip_atomic_fetch_add(cell, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    unsigned s = 0;
    for (int i = threadIdx.x; i < 512; i += 64) s += cells[i];
    sink[(size_t)blockIdx.x * 64 + threadIdx.x] = s;
  }
}

// ---------------------------------------------------------------- device binning
// BinMapper::value_to_bin (binning.rs:637-667) of one value.
__device__ __forceinline__ unsigned value_to_bin_dev(const BinFeat& bf, const double* bounds, const int* ckeys,
                                                     const int* cbins, double v) {
  if (bf.bin_type == 1) {
    if (isnan(v)) return 0u;
    // Rust `as i32`: saturating
    const int iv = v >= 2147483647.0 ? 2147483647 : v <= -2147483648.0 ? (-2147483647 - 1) : (int)v;
    if (iv < 0) return 0u;
    const int* k = ckeys + bf.off;
    int lo = 0, hi = bf.ncat;
    while (lo < hi) {
      const int m = (lo + hi) >> 1;
      if (k[m] < iv) lo = m + 1; else hi = m;
    }
    return (lo < bf.ncat && k[lo] == iv) ? (unsigned)cbins[bf.off + lo] : 0u;
  }
  if (isnan(v)) {
    if (bf.missing_type == 2) return (unsigned)(bf.num_bin - 1);
    v = 0.0;
  }
  const double* ub = bounds + bf.off;
  int l = 0, r = bf.num_bin - 1 - (bf.missing_type == 2 ? 1 : 0);
  while (l < r) {
    const int m = (r + l - 1) / 2;
    if (v <= ub[m]) r = m; else l = m + 1;
  }
  return (unsigned)l;
}

// Bins rows x ncol f32 values into rows of `stride` bytes: one thread per 4
// bytes of an output row, values loaded before use. Without `map`, byte j is
// feature j (j < F; the rest zero); with it, byte p < W holds map[p].lo in its
// low bits and, for a pair of 4-bit fields, map[p].hi in its high 4 bits
// (bytes W.. zero). STAGE: the bin bounds of every feature (nbounds doubles)
// are copied to LDS first, so the binary searches read LDS.
template <bool STAGE, bool PACK>
__global__ void __launch_bounds__(1024) k_bin_rows(const float* X, long long n, int ncol, const BinFeat* feats, int F,
                                                   const double* gbounds, int nbounds, const int* ckeys,
                                                   const int* cbins, uint8_t* out, int stride, const BinByte* map,
                                                   int W) {
  extern __shared__ double bin_lds[];
  const double* bounds = gbounds;
  if (STAGE) {
    for (int i = threadIdx.x; i < nbounds; i += blockDim.x) bin_lds[i] = gbounds[i];
    __syncthreads();
    bounds = bin_lds;
  }
  const int WW = (stride + 3) / 4;
  const long long total = n * WW;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const long long r = t / WW;
    const int w = (int)(t - r * WW);
    unsigned word = 0;
    if (PACK) {
      BinByte bb[4];
      float xl[4], xh[4];
      BinFeat fl[4], fh[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int p = 4 * w + j;
        bb[j] = p < W ? map[p] : BinByte{-1, -1};
        fl[j] = feats[max(bb[j].lo, 0)];
        fh[j] = feats[max(bb[j].hi, 0)];
        xl[j] = X[r * ncol + fl[j].col];
        xh[j] = X[r * ncol + fh[j].col];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        unsigned v = 0u;
        if (bb[j].lo >= 0) v = value_to_bin_dev(fl[j], bounds, ckeys, cbins, (double)xl[j]);
        if (bb[j].hi >= 0) v |= value_to_bin_dev(fh[j], bounds, ckeys, cbins, (double)xh[j]) << 4;
        word |= (v & 0xFFu) << (8 * j);
      }
    } else {
      float x[4];
      BinFeat bf[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int f = min(4 * w + j, F - 1);
        bf[j] = feats[f];
        x[j] = X[r * ncol + bf[j].col];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (4 * w + j < F) word |= (value_to_bin_dev(bf[j], bounds, ckeys, cbins, (double)x[j]) & 0xFFu) << (8 * j);
    }
    uint8_t* o = out + r * stride + 4 * w;
    if ((stride & 3) == 0) {
      *(unsigned*)o = word;
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (4 * w + j < stride) o[j] = (uint8_t)(word >> (8 * j));
    }
  }
}

// Column-major u8 copy [F][n] of packed row-major bins [n][S].
__global__ void k_rm_to_cm(const uint8_t* rm, long long n, int S, int F, const FeatDev* feats, uint8_t* cm) {
  const long long total = n * F;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const int f = (int)(t / n);
    const long long r = t - (long long)f * n;
    const FeatDev& d = feats[f];
    const unsigned v = rm[r * S + d.pos];
    cm[t] = (uint8_t)(d.nib < 0 ? v : (v >> (4 * d.nib)) & 0xFu);
  }
}

// Packed rows [n][S] into one u8 bin per feature [n][F].
__global__ void k_unpack_rows(const uint8_t* rm, long long n, int S, int F, const FeatDev* feats, uint8_t* out) {
  const long long total = n * F;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const long long r = t / F;
    const int f = (int)(t - r * F);
    const FeatDev& d = feats[f];
    const unsigned v = rm[r * S + d.pos];
    out[t] = (uint8_t)(d.nib < 0 ? v : (v >> (4 * d.nib)) & 0xFu);
  }
}

}  // namespace lgbmhip
