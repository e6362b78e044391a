// Evaluation metrics on the GPU (metric.rs:551-729).
//
// AUC sorts the scores once (a radix sort of order-preserving 64-bit keys),
// then sums over runs of equal scores from prefix counts. Unweighted, every
// partial sum is an integer or half-integer below 2^53, so the value is the
// reference's bit for bit whatever the order of ties. Pointwise losses sum in
// a fixed two-level order (deterministic, not the reference's row order).
// NDCG selects each query's top documents in the reference's order (score
// descending, index ascending on ties) and reproduces its per-query DCG
// chains and its per-thread partial sums exactly.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "engine.h"
#include "metrics.h"

namespace lgbmhip {

#define MCHECK(x)                                                                              \
  do {                                                                                         \
    hipError_t e_ = (x);                                                                       \
    if (e_ != hipSuccess) throw Error(LGBMHIP_INTERNAL, std::string("HIP: ") + hipGetErrorString(e_)); \
  } while (0)

namespace {

struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
  void* get(size_t want) {
    if (want > bytes) {
      if (p) (void)hipFree(p);
      p = nullptr;
      bytes = 0;
      MCHECK(hipMalloc(&p, std::max<size_t>(want, 256)));
      bytes = std::max<size_t>(want, 256);
    }
    return p;
  }
  ~Buf() {
    if (p) (void)hipFree(p);
  }
};

constexpr int kRedBlocks = 1024;   // blocks of the fixed-order reductions
constexpr int kRedThreads = 256;

// Sum of one value per thread in a fixed tree order; the result is valid in thread 0.
template <class T>
__device__ T block_sum(T v, T* sm) {
  const int t = threadIdx.x;
  sm[t] = v;
  __syncthreads();
  for (int w = kRedThreads / 2; w > 0; w >>= 1) {
    if (t < w) sm[t] = sm[t] + sm[t + w];
    __syncthreads();
  }
  return sm[0];
}

// ---------------------------------------------------------------- AUC
// Key whose ascending order is descending score; -0.0 and +0.0 share one key
// (the reference compares scores with !=).
__device__ __forceinline__ unsigned long long desc_key(double s) {
  unsigned long long u = s == 0.0 ? 0ull : (unsigned long long)__double_as_longlong(s);
  const unsigned long long asc = (u >> 63) ? ~u : (u | (1ull << 63));
  return ~asc;
}

// Sort input: keys, and the row (weighted) or the positive flag (unweighted,
// so the sort carries the label and no gather follows).
template <bool W>
__global__ void k_auc_keys(const double* score, const float* label, long long n, unsigned long long* keys, int* rows,
                           uint8_t* flags) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    keys[i] = desc_key(score[i]);
    if (W) rows[i] = (int)i;
    else flags[i] = label[i] > 0.0f ? 1 : 0;
  }
}

// Per sorted position: the position itself where a run of equal keys starts
// (0 elsewhere) and, weighted, the positive / negative weights.
template <bool W>
__global__ void k_auc_marks(const unsigned long long* keys, const int* rows, const float* label, const float* weight,
                            long long n, int* head, double* wpos, double* wneg) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    head[i] = (i == 0 || keys[i] != keys[i - 1]) ? (int)i : 0;
    if (W) {
      const int r = rows[i];
      const float l = label[r];
      const double w = (double)weight[r];
      wneg[i] = (double)(l <= 0.0f ? 1 : 0) * w;
      wpos[i] = (double)(l > 0.0f ? 1 : 0) * w;
    }
  }
}

// One term per run of equal scores [a, b), taken at its last position:
// unweighted, 2 x neg (pos / 2 + positives before a) = neg (P_a + P_b) with
// P_x the positives in [0, x); weighted, neg_w (pos_w / 2 + before_w).
template <bool W>
__global__ void __launch_bounds__(kRedThreads) k_auc_runs(const unsigned long long* keys, long long n, const int* pre,
                                                         const uint8_t* pos, const int* head_max, const double* pwpre,
                                                         const double* nwpre, const double* wpos, const double* wneg,
                                                         long long* ipart, double* dpart) {
  __shared__ long long ism[kRedThreads];
  __shared__ double dsm[kRedThreads];
  long long iacc = 0;
  double dacc = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    if (i + 1 < n && keys[i + 1] == keys[i]) continue;
    const int a = head_max[i];
    if (W) {
      const double pa = pwpre[a];
      const double pos_w = (pwpre[i] + wpos[i]) - pa;
      const double neg_w = (nwpre[i] + wneg[i]) - nwpre[a];
      dacc += neg_w * (pos_w * 0.5 + pa);
    } else {
      const long long pa = pre[a], pb = (long long)pre[i] + pos[i];
      const long long neg = (i + 1 - a) - (pb - pa);
      iacc += neg * (pa + pb);
    }
  }
  if (W) {
    const double v = block_sum(dacc, dsm);
    if (threadIdx.x == 0) dpart[blockIdx.x] = v;
  } else {
    const long long v = block_sum(iacc, ism);
    if (threadIdx.x == 0) ipart[blockIdx.x] = v;
  }
}

// metric.rs:724-728: accum / (sum_pos (sum_weights - sum_pos)), or 1.
template <bool W>
__global__ void __launch_bounds__(kRedThreads) k_auc_final(const long long* ipart, const double* dpart, int nparts,
                                                          const int* pre, const uint8_t* pos, const double* pwpre,
                                                          const double* wpos, long long n, double sum_weights,
                                                          double* out) {
  __shared__ long long ism[kRedThreads];
  __shared__ double dsm[kRedThreads];
  long long iacc = 0;
  double dacc = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
    if (W) dacc += dpart[i];
    else iacc += ipart[i];
  }
  double accum, sum_pos;
  if (W) {
    accum = block_sum(dacc, dsm);
    sum_pos = pwpre[n - 1] + wpos[n - 1];
  } else {
    accum = (double)block_sum(iacc, ism) * 0.5;
    sum_pos = (double)((long long)pre[n - 1] + pos[n - 1]);
  }
  if (threadIdx.x == 0)
    out[0] = (sum_pos > 0.0 && sum_pos != sum_weights) ? accum / (sum_pos * (sum_weights - sum_pos)) : 1.0;
}

struct MaxOp {
  __device__ __host__ int operator()(int a, int b) const { return a > b ? a : b; }
};

// ---------------------------------------------------------------- pointwise
// metric.rs:602-625 (binary, sigmoid transform binary.rs:164-166) and
// metric.rs:629-653 with multiclass_loss :133-153 (softmax multiclass.rs:16-30).
template <int KIND>
__global__ void __launch_bounds__(kRedThreads) k_metric_point(const double* score, const float* label,
                                                             const float* weight, long long n, int K, double sigmoid,
                                                             int top_k, double nle, double* part) {
  __shared__ double sm[kRedThreads];
  constexpr double kEps = (double)1e-15f;
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float lbl = label[i];
    double l;
    if (KIND == kBinaryLogloss || KIND == kBinaryError) {
      const double p = 1.0 / (1.0 + exp(-sigmoid * score[i]));
      if (KIND == kBinaryLogloss) {
        if (lbl <= 0.0f) l = 1.0 - p > kEps ? -log(1.0 - p) : nle;
        else l = p > kEps ? -log(p) : nle;
      } else {
        l = p <= 0.5 ? (double)(lbl > 0.0f ? 1 : 0) : (double)(lbl <= 0.0f ? 1 : 0);
      }
    } else {
      const int k = (int)lbl;
      double wmax = score[i];
      for (int c = 1; c < K; ++c) {
        const double v = score[(size_t)c * n + i];
        wmax = v < wmax ? wmax : v;
      }
      double wsum = 0.0;
      for (int c = 0; c < K; ++c) wsum += exp(score[(size_t)c * n + i] - wmax);
      const double rk = exp(score[(size_t)k * n + i] - wmax) / wsum;
      if (KIND == kMultiLogloss) {
        l = rk > kEps ? -log(rk) : nle;
      } else {
        int larger = 0;
        l = 0.0;
        for (int c = 0; c < K; ++c) {
          const double rc = exp(score[(size_t)c * n + i] - wmax) / wsum;
          if (rc >= rk) ++larger;
          if (larger > top_k) {
            l = 1.0;
            break;
          }
        }
      }
    }
    acc += weight ? l * (double)weight[i] : l;
  }
  const double v = block_sum(acc, sm);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ void __launch_bounds__(kRedThreads) k_metric_final(const double* part, int nparts, double sum_weights,
                                                             double* out) {
  __shared__ double sm[kRedThreads];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) acc += part[i];
  const double v = block_sum(acc, sm);
  if (threadIdx.x == 0) out[0] = v / sum_weights;
}

// ---------------------------------------------------------------- NDCG
constexpr int kNdcgStage = 4096;   // query scores staged in LDS up to this many documents
constexpr int kMaxGain = 128;
constexpr int kMaxEvalAt = 32;

struct Cand {
  double s;
  int i;
};
// a ranks before b: higher score, or equal score and lower index (dcg.rs:167-179)
__device__ __forceinline__ bool before(double sa, int ia, double sb, int ib) {
  return sa > sb || (!(sb > sa) && ia < ib);
}

// One wave per query: DCG at each cut-off over the documents selected one
// position at a time in the reference's order, the max DCG from label
// counts (dcg.rs:88-103), and the query's term tmp x inverse max DCG, or 1
// for a query without relevant documents (metric.rs:564-576).
__global__ void __launch_bounds__(64) k_ndcg_query(const double* score, const float* label, const int* qb, int nq,
                                                   const int* eval_at, int nk, const double* gain, int ngain,
                                                   const double* disc, double* terms) {
  __shared__ double ss[kNdcgStage];
  __shared__ int lc[kMaxGain];
  __shared__ double tmp[kMaxEvalAt];
  const int q = blockIdx.x, lane = threadIdx.x;
  const int s0 = qb[q], cnt = qb[q + 1] - s0;
  const double* sq = score + s0;
  const float* lq = label + s0;
  const bool staged = cnt <= kNdcgStage;
  for (int i = lane; i < ngain; i += 64) lc[i] = 0;
  __syncthreads();
  for (int i = lane; i < cnt; i += 64) {
    if (staged) ss[i] = sq[i];
    atomicAdd(&lc[(int)lq[i]], 1);
  }
  __syncthreads();
  const double* sv = staged ? ss : sq;
  const int kmax = min(eval_at[nk - 1], cnt);
  // DCG (dcg.rs:105-118): position j holds the j-th document in order
  double cur = 0.0;
  int m = 0;
  while (m < nk && min(eval_at[m], cnt) == 0) tmp[m++] = 0.0;
  double ps = 0.0;
  int pi = -1;
  for (int j = 0; j < kmax; ++j) {
    double bs = 0.0;
    int bi = -1;
    for (int i = lane; i < cnt; i += 64) {
      const double si = sv[i];
      const bool after = pi < 0 || before(ps, pi, si, i);
      if (after && (bi < 0 || before(si, i, bs, bi))) {
        bs = si;
        bi = i;
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double os = __shfl_xor(bs, o);
      const int oi = __shfl_xor(bi, o);
      if (oi >= 0 && (bi < 0 || before(os, oi, bs, bi))) {
        bs = os;
        bi = oi;
      }
    }
    // lane 0's choice for every lane (one order even for NaN scores)
    bs = __shfl(bs, 0);
    bi = __shfl(bi, 0);
    if (bi < 0) break;
    ps = bs;
    pi = bi;
    cur += gain[(int)lq[bi]] * disc[j];
    while (m < nk && min(eval_at[m], cnt) == j + 1) tmp[m++] = cur;
  }
  while (m < nk) tmp[m++] = cur;
  if (lane != 0) return;
  // max DCG (dcg.rs:88-103) and the query's terms (metric.rs:564-576)
  double mcur = 0.0, inv0 = 0.0;
  int top = ngain - 1, left = 0;
  for (int mm = 0; mm < nk; ++mm) {
    const int k = min(eval_at[mm], cnt);
    for (int j = left; j < k; ++j) {
      while (top > 0 && lc[top] <= 0) --top;
      mcur += disc[j] * gain[top];
      lc[top] -= 1;
    }
    left = k;
    const double inv = mcur > 0.0 ? 1.0 / mcur : -1.0;
    if (mm == 0) inv0 = inv;
    terms[(size_t)q * nk + mm] = inv0 <= 0.0 ? 1.0 : tmp[mm] * inv;
  }
}

// metric.rs:556-599: libgomp static chunks of queries, one sequential sum per
// chunk, chunk sums added in chunk order, divided by the query count.
__global__ void __launch_bounds__(256) k_ndcg_sum(const double* terms, int nq, int nk, int threads, double* bufs,
                                                  double sum_qw, double* out) {
  const int qc = nq / threads, r = nq % threads;
  for (int t = threadIdx.x; t < threads; t += blockDim.x) {
    const int start = t * qc + min(t, r), len = qc + (t < r ? 1 : 0);
    double acc[kMaxEvalAt];
    for (int m = 0; m < nk; ++m) acc[m] = 0.0;
    for (int q = start; q < start + len; ++q)
      for (int m = 0; m < nk; ++m) acc[m] += terms[(size_t)q * nk + m];
    for (int m = 0; m < nk; ++m) bufs[(size_t)t * nk + m] = acc[m];
  }
  __syncthreads();
  if (threadIdx.x < nk) {
    const int m = threadIdx.x;
    double v = 0.0;
    for (int t = 0; t < threads; ++t) v += bufs[(size_t)t * nk + m];
    out[m] = v / sum_qw;
  }
}

}  // namespace

struct MetricWs {
  Buf keys0, keys1, rows0, rows1, pre, head, headmax, wpos, wneg, pwpre, nwpre, part, temp, res;
  Buf terms, bufs, qtab;
};

MetricWs* metric_ws_new() { return new MetricWs(); }
void metric_ws_free(MetricWs* ws) { delete ws; }

static int red_grid(int64_t n) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(kRedBlocks, (n + kRedThreads - 1) / kRedThreads));
}

double metric_auc(MetricWs* ws, hipStream_t s, const double* score, const float* label, const float* weight,
                  int64_t n, double sum_weights) {
  if (n <= 0) throw Error(LGBMHIP_INVALID_DATA, "AUC of no rows");
  if (n > INT32_MAX) throw Error(LGBMHIP_UNSUPPORTED, "AUC over more than 2^31-1 rows");
  const bool W = weight != nullptr;
  auto* k0 = (unsigned long long*)ws->keys0.get(n * 8);
  auto* k1 = (unsigned long long*)ws->keys1.get(n * 8);
  auto* head = (int*)ws->head.get(n * 4);
  auto* hmax = (int*)ws->headmax.get(n * 4);
  int *r0 = nullptr, *r1 = nullptr, *pre = nullptr;
  uint8_t *f0 = nullptr, *f1 = nullptr;
  double *wpos = nullptr, *wneg = nullptr, *pwpre = nullptr, *nwpre = nullptr;
  if (W) {
    r0 = (int*)ws->rows0.get(n * 4);
    r1 = (int*)ws->rows1.get(n * 4);
    wpos = (double*)ws->wpos.get(n * 8);
    wneg = (double*)ws->wneg.get(n * 8);
    pwpre = (double*)ws->pwpre.get(n * 8);
    nwpre = (double*)ws->nwpre.get(n * 8);
  } else {
    f0 = (uint8_t*)ws->rows0.get(n);
    f1 = (uint8_t*)ws->rows1.get(n);
    pre = (int*)ws->pre.get(n * 4);
  }
  const int g = red_grid(n);
  auto* ipart = (long long*)ws->part.get(kRedBlocks * 8);
  auto* dpart = (double*)ipart;
  auto* res = (double*)ws->res.get(8);
  size_t tb = 0, tb2 = 0, tb3 = 0;
  if (W) {
    hipLaunchKernelGGL(k_auc_keys<true>, dim3(g), dim3(kRedThreads), 0, s, score, label, (long long)n, k0, r0, f0);
    MCHECK(rocprim::radix_sort_pairs(nullptr, tb, k0, k1, r0, r1, (unsigned)n, 0, 64, s));
    MCHECK(rocprim::deterministic_exclusive_scan(nullptr, tb2, wpos, pwpre, 0.0, (size_t)n, rocprim::plus<double>(), s));
  } else {
    hipLaunchKernelGGL(k_auc_keys<false>, dim3(g), dim3(kRedThreads), 0, s, score, label, (long long)n, k0, r0, f0);
    MCHECK(rocprim::radix_sort_pairs(nullptr, tb, k0, k1, f0, f1, (unsigned)n, 0, 64, s));
    MCHECK(rocprim::exclusive_scan(nullptr, tb2, f1, pre, 0, (size_t)n, rocprim::plus<int>(), s));
  }
  MCHECK(rocprim::inclusive_scan(nullptr, tb3, head, hmax, (size_t)n, MaxOp(), s));
  void* temp = ws->temp.get(std::max(tb, std::max(tb2, tb3)));
  if (W) {
    MCHECK(rocprim::radix_sort_pairs(temp, tb, k0, k1, r0, r1, (unsigned)n, 0, 64, s));
    hipLaunchKernelGGL(k_auc_marks<true>, dim3(g), dim3(kRedThreads), 0, s, k1, r1, label, weight, (long long)n,
                       head, wpos, wneg);
  } else {
    MCHECK(rocprim::radix_sort_pairs(temp, tb, k0, k1, f0, f1, (unsigned)n, 0, 64, s));
    hipLaunchKernelGGL(k_auc_marks<false>, dim3(g), dim3(kRedThreads), 0, s, k1, r1, label, weight, (long long)n,
                       head, wpos, wneg);
  }
  MCHECK(rocprim::inclusive_scan(temp, tb3, head, hmax, (size_t)n, MaxOp(), s));
  if (W) {
    MCHECK(rocprim::deterministic_exclusive_scan(temp, tb2, wpos, pwpre, 0.0, (size_t)n, rocprim::plus<double>(), s));
    MCHECK(rocprim::deterministic_exclusive_scan(temp, tb2, wneg, nwpre, 0.0, (size_t)n, rocprim::plus<double>(), s));
    hipLaunchKernelGGL(k_auc_runs<true>, dim3(g), dim3(kRedThreads), 0, s, k1, (long long)n, pre, f1, hmax, pwpre,
                       nwpre, wpos, wneg, ipart, dpart);
    hipLaunchKernelGGL(k_auc_final<true>, dim3(1), dim3(kRedThreads), 0, s, ipart, dpart, g, pre, f1, pwpre, wpos,
                       (long long)n, sum_weights, res);
  } else {
    MCHECK(rocprim::exclusive_scan(temp, tb2, f1, pre, 0, (size_t)n, rocprim::plus<int>(), s));
    hipLaunchKernelGGL(k_auc_runs<false>, dim3(g), dim3(kRedThreads), 0, s, k1, (long long)n, pre, f1, hmax, pwpre,
                       nwpre, wpos, wneg, ipart, dpart);
    hipLaunchKernelGGL(k_auc_final<false>, dim3(1), dim3(kRedThreads), 0, s, ipart, dpart, g, pre, f1, pwpre, wpos,
                       (long long)n, sum_weights, res);
  }
  MCHECK(hipGetLastError());
  double out = 0.0;
  MCHECK(hipMemcpyAsync(&out, res, 8, hipMemcpyDeviceToHost, s));
  MCHECK(hipStreamSynchronize(s));
  return out;
}

double metric_point(MetricWs* ws, hipStream_t s, int kind, const double* score, const float* label,
                    const float* weight, int64_t n, int K, double sigmoid, int top_k, double neg_log_eps,
                    double sum_weights) {
  if (n <= 0) throw Error(LGBMHIP_INVALID_DATA, "metric over no rows");
  const int g = red_grid(n);
  auto* part = (double*)ws->part.get(kRedBlocks * 8);
  auto* res = (double*)ws->res.get(8);
  const dim3 grid(g), blk(kRedThreads);
  switch (kind) {
    case kBinaryLogloss:
      hipLaunchKernelGGL(k_metric_point<kBinaryLogloss>, grid, blk, 0, s, score, label, weight, (long long)n, K,
                         sigmoid, top_k, neg_log_eps, part);
      break;
    case kBinaryError:
      hipLaunchKernelGGL(k_metric_point<kBinaryError>, grid, blk, 0, s, score, label, weight, (long long)n, K,
                         sigmoid, top_k, neg_log_eps, part);
      break;
    case kMultiLogloss:
      hipLaunchKernelGGL(k_metric_point<kMultiLogloss>, grid, blk, 0, s, score, label, weight, (long long)n, K,
                         sigmoid, top_k, neg_log_eps, part);
      break;
    case kMultiError:
      hipLaunchKernelGGL(k_metric_point<kMultiError>, grid, blk, 0, s, score, label, weight, (long long)n, K, sigmoid,
                         top_k, neg_log_eps, part);
      break;
    default:
      throw Error(LGBMHIP_INVALID_PARAMETER, "unknown pointwise metric");
  }
  hipLaunchKernelGGL(k_metric_final, dim3(1), blk, 0, s, part, g, sum_weights, res);
  MCHECK(hipGetLastError());
  double out = 0.0;
  MCHECK(hipMemcpyAsync(&out, res, 8, hipMemcpyDeviceToHost, s));
  MCHECK(hipStreamSynchronize(s));
  return out;
}

void metric_ndcg(MetricWs* ws, hipStream_t s, const double* score, const float* label, const int* qb, int nq,
                 const int* eval_at_host, int nk, const double* gain_host, int ngain, const double* disc_host,
                 int ndisc, int threads, double* out_host) {
  if (nq <= 0) throw Error(LGBMHIP_INVALID_DATA, "NDCG of no queries");
  if (nk <= 0 || nk > kMaxEvalAt) throw Error(LGBMHIP_INVALID_PARAMETER, "1 to 32 NDCG cut-offs");
  for (int m = 0; m < nk; ++m)
    if (eval_at_host[m] <= 0 || eval_at_host[m] > ndisc || (m > 0 && eval_at_host[m] < eval_at_host[m - 1]))
      throw Error(LGBMHIP_INVALID_PARAMETER, "NDCG cut-offs must ascend within [1, 10000]");
  if (ngain <= 0 || ngain > kMaxGain) throw Error(LGBMHIP_UNSUPPORTED, "more than 128 label gains");
  threads = std::max(1, threads);
  // tables: eval_at (int) | gain | discount
  const size_t off_g = 256, off_d = off_g + (size_t)ngain * 8;
  auto* tab = (unsigned char*)ws->qtab.get(off_d + (size_t)ndisc * 8);
  MCHECK(hipMemcpyAsync(tab, eval_at_host, (size_t)nk * 4, hipMemcpyHostToDevice, s));
  MCHECK(hipMemcpyAsync(tab + off_g, gain_host, (size_t)ngain * 8, hipMemcpyHostToDevice, s));
  MCHECK(hipMemcpyAsync(tab + off_d, disc_host, (size_t)ndisc * 8, hipMemcpyHostToDevice, s));
  auto* terms = (double*)ws->terms.get((size_t)nq * nk * 8);
  auto* bufs = (double*)ws->bufs.get((size_t)threads * nk * 8);
  auto* res = (double*)ws->res.get((size_t)nk * 8);
  hipLaunchKernelGGL(k_ndcg_query, dim3(nq), dim3(64), 0, s, score, label, qb, nq, (const int*)tab, nk,
                     (const double*)(tab + off_g), ngain, (const double*)(tab + off_d), terms);
  hipLaunchKernelGGL(k_ndcg_sum, dim3(1), dim3(256), 0, s, terms, nq, nk, threads, bufs, (double)nq, res);
  MCHECK(hipGetLastError());
  MCHECK(hipMemcpyAsync(out_host, res, (size_t)nk * 8, hipMemcpyDeviceToHost, s));
  MCHECK(hipStreamSynchronize(s));
}

}  // namespace lgbmhip
