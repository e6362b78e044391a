// Evaluation metrics over device-resident scores (metric.rs:551-729): the
// launchers the context calls; every pointer argument is device memory unless
// its name says host.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lgbmhip {

// Device workspace of the metric launchers, grown on demand and reused.
struct MetricWs;
MetricWs* metric_ws_new();
void metric_ws_free(MetricWs* ws);

// AUC of class-0 scores (metric.rs:701-729). weight may be null; sum_weights
// is the metric's weight total (rows when unweighted). Returns the value.
double metric_auc(MetricWs* ws, hipStream_t s, const double* score, const float* label, const float* weight,
                  int64_t n, double sum_weights);

// Pointwise losses (metric.rs:602-653) averaged over the weight total.
enum PointKind { kBinaryLogloss = 0, kBinaryError = 1, kMultiLogloss = 2, kMultiError = 3 };
double metric_point(MetricWs* ws, hipStream_t s, int kind, const double* score, const float* label,
                    const float* weight, int64_t n, int K, double sigmoid, int top_k, double neg_log_eps,
                    double sum_weights);

// NDCG at each of the nk ascending cut-offs (metric.rs:551-600, dcg.rs:68-121):
// per-query terms, then the per-thread partial sums of `threads` static
// chunks of queries, summed in thread order and divided by the query count.
// gain_host / disc_host are the label gains and position discounts.
void metric_ndcg(MetricWs* ws, hipStream_t s, const double* score, const float* label, const int* qb, int nq,
                 const int* eval_at_host, int nk, const double* gain_host, int ngain, const double* disc_host,
                 int ndisc, int threads, double* out_host);

}  // namespace lgbmhip
