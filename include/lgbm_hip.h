/*
 * lgbm_hip.h — C ABI of the MI355X GBDT engine (liblgbm_hip.so).
 *
 * The drop-in boundary for lightgbm-rust's training/prediction hot path.
 * Every entry point is `extern "C"`, takes plain pointers and sizes, borrows
 * caller memory only for the duration of the call, catches every C++
 * exception and GPU error, and returns an lgbmhip_status. A context owns its
 * device memory and one HIP stream; it is not thread-safe (one caller at a
 * time, as `&mut Gbdt` in the reference).
 *
 * Reference interfaces each entry point stands in for (paths relative to
 * crates/lgbm-core/src/ of zrluety/lightgbm-rust):
 *   lgbmhip_status codes      <- LgbmError variants            error.rs:6-18
 *   lgbmhip_bin_dense         <- Dataset::from_matrix (binning) dataset.rs:360-422
 *   lgbmhip_find_mappers      <- BinMapper::find_bin (sample)   dataset.rs:393-397, binning.rs:398-517
 *   lgbmhip_upload_raw        <- Dataset::from_matrix push pass dataset.rs:594-663, binning.rs:637-667
 *   lgbmhip_bin_rows_device   <- Dataset::from_matrix_with_reference (bins) dataset.rs:456-520
 *   lgbmhip_upload_bins       <- SerialTreeLearner::new         learner/mod.rs:220-310
 *   lgbmhip_set_metadata      <- Objective::init / Metadata     objective/mod.rs:43-102, dataset.rs:104-130
 *   lgbmhip_set_split_params  <- SplitParams                    learner/split.rs:33-46
 *   lgbmhip_set_train_opts    <- Gbdt::new                      boosting.rs:206-337
 *   lgbmhip_train_iteration   <- Gbdt::train_one_iter           boosting.rs:957-1139
 *   lgbmhip_train_tree        <- SerialTreeLearner::train       learner/mod.rs:572-695
 *   lgbmhip_tree_*            <- Tree arrays / SplitArgs        tree.rs:120-178, 254-297
 *   lgbmhip_copy_scores       <- Gbdt::train_scores             boosting.rs:339-341
 *   lgbmhip_add_leaf_outputs  <- SerialTreeLearner::add_leaf_outputs learner/mod.rs:537, partition.rs:178-191
 *   lgbmhip_gradients         <- RowObjective::gradients        objective/mod.rs:52, binary.rs:101-127,
 *                                                               multiclass.rs:126-154, rank.rs:63-108
 *   lgbmhip_discretize        <- GradientDiscretizer::discretize learner/quant.rs:92-129
 *   lgbmhip_build_histogram*  <- MultiValBin::construct(_int)   multi_val_bin.rs:556-611, 710-753
 *   lgbmhip_find_best_split*  <- find_best_threshold(_int)      learner/split.rs:246-271, 930-950
 *   lgbmhip_partition         <- DataPartition::split           learner/partition.rs:107-174
 *   lgbmhip_predict_raw       <- Gbdt::predict_matrix_early_stop predict.rs:231-339
 *   lgbmhip_predict_binned    <- Tree::add_prediction_to_score  tree.rs:605-652
 *   lgbmhip_add_valid         <- Gbdt::add_valid                boosting.rs:378-411
 *   lgbmhip_copy_valid_scores <- Gbdt::valid_scores             boosting.rs:1282-1284
 *   lgbmhip_eval              <- Metric::eval                   metric.rs:525-545, 551-729
 *   lgbmhip_set_predict_window <- predict_window / prediction_early_stop predict.rs:178-202
 *   lgbmhip_predict_leaf      <- PredictKind::LeafIndex         predict.rs:272-279
 *   lgbmhip_rollback_one_iter <- Gbdt::rollback_one_iter        boosting.rs:1215-1239
 *   lgbmhip_save_model        <- GBDT::SaveModelToString        io.rs:13-88, tree.rs:718-770
 *   lgbmhip_load_model        <- GBDT::LoadModelFromString      io.rs:156-343
 *   lgbmhip_check_model       <- LoadModelFromString / Tree::from_model_str checks io.rs:156-343, tree.rs:869-1047
 *   lgbmhip_bins_feature_infos <- Dataset::feature_infos        dataset.rs:770-776
 */
#ifndef LGBM_HIP_H_
#define LGBM_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t lgbmhip_status;
#define LGBMHIP_OK 0
#define LGBMHIP_INVALID_PARAMETER 1
#define LGBMHIP_INVALID_DATA 2
#define LGBMHIP_MODEL_FORMAT 3
#define LGBMHIP_UNSUPPORTED 4
#define LGBMHIP_INTERNAL 5

typedef struct lgbmhip_ctx lgbmhip_ctx;
typedef struct lgbmhip_bins lgbmhip_bins;

/* Parameters of host binning (Config fields read by dataset.rs:860-910). */
typedef struct {
  int32_t max_bin;
  int32_t min_data_in_bin;
  int32_t min_data_in_leaf;
  int32_t bin_construct_sample_cnt;
  int32_t data_random_seed;
  int32_t feature_pre_filter;
  int32_t use_missing;
  int32_t zero_as_missing;
  int32_t num_threads; /* host threads for binning; <= 0: all */
} lgbmhip_bin_params;

/* One used feature as the learner sees it (split.rs:18-29 FeatureMeta). */
typedef struct {
  int32_t real_index;    /* column of the raw matrix */
  int32_t num_bin;
  int32_t missing_type;  /* 0 none, 1 zero, 2 NaN */
  int32_t default_bin;
  int32_t most_freq_bin;
  int32_t bin_type;      /* 0 numerical, 1 categorical */
} lgbmhip_feature_meta;

/* split.rs:33-46 plus the tree-shape limits of the learner. */
typedef struct {
  double lambda_l1, lambda_l2, max_delta_step, path_smooth;
  double min_sum_hessian_in_leaf, min_gain_to_split, cat_l2, cat_smooth;
  int32_t min_data_in_leaf, max_cat_to_onehot, max_cat_threshold, min_data_per_group;
  int32_t num_leaves, max_depth;
} lgbmhip_split_params;

/* Boosting options (the Config fields Gbdt::new reads for this path). */
typedef struct {
  int32_t objective;            /* 0 binary, 1 multiclass (softmax), 2 lambdarank */
  int32_t num_class;
  double sigmoid;
  double learning_rate;
  int32_t boost_from_average;
  int32_t use_quantized_grad;
  int32_t num_grad_quant_bins;
  int32_t stochastic_rounding;
  int32_t quant_train_renew_leaf;
  int32_t seed;                 /* Config::seed (quantizer generators) */
  int32_t quant_table_threads;  /* thread count the random tables are drawn for (reference num_threads) */
  int32_t lambdarank_truncation_level;
  int32_t lambdarank_norm;
  int32_t use_graph;            /* capture each tree's launches in a hipGraph */
  double bagging_fraction;      /* row bagging (sample_strategy.rs): on when < 1 and bagging_freq > 0 */
  int32_t bagging_freq;
  int32_t bagging_seed;
  double pos_bagging_fraction;  /* balanced bagging (binary objective): fractions of rows with label > 0 / <= 0 */
  double neg_bagging_fraction;
  int32_t data_sample_strategy; /* 0 bagging, 1 goss */
  int32_t sample_threads;       /* GOSS: the reference's num_threads, which sets its row blocks (>= 1) */
  double top_rate;              /* GOSS: share of rows kept by |g h| and share sampled from the rest */
  double other_rate;
  int32_t bagging_by_query;     /* row bagging draws whole queries (query boundaries required) */
  int32_t batch_trees;          /* num_class > 1: class trees per shared growth launch; 1: chosen by the engine,
                                   0: one tree per launch */
} lgbmhip_train_opts;

/* One metric of lgbmhip_eval (metric.rs:17-43 MetricKind, the Config fields it reads). */
#define LGBMHIP_METRIC_BINARY_LOGLOSS 0
#define LGBMHIP_METRIC_BINARY_ERROR 1
#define LGBMHIP_METRIC_AUC 2
#define LGBMHIP_METRIC_MULTI_LOGLOSS 3
#define LGBMHIP_METRIC_MULTI_ERROR 4
#define LGBMHIP_METRIC_NDCG 5
typedef struct {
  int32_t metric;              /* LGBMHIP_METRIC_* */
  int32_t multi_error_top_k;   /* multi_error: a row counts as an error above this many classes scoring >= its own */
  int32_t num_threads;         /* NDCG: the reference's per-thread partial sums over queries (1: query order) */
  int32_t num_eval_at;         /* NDCG cut-offs, ascending; 0: 1..5 */
  const int32_t* eval_at;
  int32_t num_label_gain;      /* 0: 2^i - 1 for i < 31 */
  const double* label_gain;
} lgbmhip_metric_opts;

const char* lgbmhip_last_error(const lgbmhip_ctx* ctx);
const char* lgbmhip_global_last_error(void);

/* ---- host binning (BinMapper stays on the host) ---- */
lgbmhip_status lgbmhip_bin_dense(const float* X, int64_t n, int32_t ncol, const int32_t* categorical, int32_t n_cat,
                                 const lgbmhip_bin_params* params, lgbmhip_bins** out);
void lgbmhip_bins_free(lgbmhip_bins* b);
int32_t lgbmhip_bins_num_features(const lgbmhip_bins* b);
/* The number of feature groups the reference's exclusive feature bundling
 * forms over the used features (replaces Dataset::num_feature_groups,
 * crates/lgbm-core/src/dataset.rs:721-723, with the groups of
 * feature_groups.rs:169-211). The engine stores every feature on its own;
 * the count feeds the bagging subset-mode rule (sample_strategy.rs:198-201). */
int32_t lgbmhip_bins_num_groups(const lgbmhip_bins* b);
/* metas: num_features entries; upper: per feature num_bin values concatenated (bin upper bounds, or the category of each bin). */
lgbmhip_status lgbmhip_bins_metas(const lgbmhip_bins* b, lgbmhip_feature_meta* metas, double* upper);
int64_t lgbmhip_bins_num_upper(const lgbmhip_bins* b);
/* The binned matrix, n x num_features row-major u8 (num_bin <= 256 for every feature). */
lgbmhip_status lgbmhip_bins_matrix(const lgbmhip_bins* b, uint8_t* out);
/* Bin raw values with the mappers of `b` (validation / prediction data). */
lgbmhip_status lgbmhip_bins_apply(const lgbmhip_bins* b, const float* X, int64_t n, int32_t ncol, uint8_t* out);

/* The mappers of lgbmhip_bin_dense without the bin matrix (the sample-based
 * BinMapper search stays on the host); lgbmhip_upload_raw then bins the rows
 * on the device. lgbmhip_bins_matrix fails on such a handle. */
lgbmhip_status lgbmhip_find_mappers(const float* X, int64_t n, int32_t ncol, const int32_t* categorical,
                                    int32_t n_cat, const lgbmhip_bin_params* params, lgbmhip_bins** out);

/* ---- context ---- */
lgbmhip_status lgbmhip_create(int32_t device, lgbmhip_ctx** out);
void lgbmhip_destroy(lgbmhip_ctx* ctx);

/* How the next upload stores its rows in HBM: u4 = 1 (the default) keeps
 * the features of at most 16 bins as 4-bit fields, two per byte after the
 * byte-wide features; u4 = 0 keeps every feature a byte. Replaces the
 * reference's storage choice BinKind::dense (bin.rs:314-324), whose Dense4
 * (bin.rs:92-121) packs the same features. Trees do not depend on it. */
lgbmhip_status lgbmhip_set_bin_packing(lgbmhip_ctx* ctx, int32_t u4);
/* The reference's feature-group count of the uploaded data, for callers of
 * lgbmhip_upload_bins (lgbmhip_upload_raw takes it from its mappers): the
 * value of lgbmhip_bins_num_groups, or the number of used features when the
 * reference would not bundle (enable_bundle false); 0 resets to the latter.
 * Replaces the count Dataset::num_feature_groups (dataset.rs:721-723) hands
 * the bagging subset-mode rule (sample_strategy.rs:198-201). */
lgbmhip_status lgbmhip_set_feature_groups(lgbmhip_ctx* ctx, int32_t num_groups);
/* The uploaded rows: out3 = {bytes per packed row, row stride in bytes,
 * features stored as 4-bit fields}. No counterpart in the reference
 * (Dataset's storage sizes): for reports and tests. */
lgbmhip_status lgbmhip_bin_layout(lgbmhip_ctx* ctx, int32_t* out3);

/* Uploads the binned matrix once (row-major and feature-major copies in HBM). */
lgbmhip_status lgbmhip_upload_bins(lgbmhip_ctx* ctx, const lgbmhip_feature_meta* metas, int32_t num_features,
                                   const double* upper, const uint8_t* bins, int64_t num_rows, int32_t num_total_features);
/* lgbmhip_upload_bins with the binning pass on the device: rows x ncol f32
 * values (host memory, or device memory when x_on_device = 1) binned with
 * the mappers of `mappers` (value_to_bin) straight into HBM. */
lgbmhip_status lgbmhip_upload_raw(lgbmhip_ctx* ctx, const lgbmhip_bins* mappers, const float* X, int64_t rows,
                                  int32_t ncol, int32_t x_on_device);
/* Device rows (rows x ncol f32) binned with the uploaded training data's
 * mappers into device u8 [rows][F] (the input of binned prediction). */
lgbmhip_status lgbmhip_bin_rows_device(lgbmhip_ctx* ctx, const float* d_X, int64_t rows, int32_t ncol,
                                       uint8_t* d_out);
lgbmhip_status lgbmhip_set_metadata(lgbmhip_ctx* ctx, const float* label, const float* weight,
                                    const int32_t* query_boundaries, int32_t num_queries);
lgbmhip_status lgbmhip_set_split_params(lgbmhip_ctx* ctx, const lgbmhip_split_params* p);
lgbmhip_status lgbmhip_set_train_opts(lgbmhip_ctx* ctx, const lgbmhip_train_opts* o);

/* One boosting iteration: gradients for all K classes, K trees, score update.
 * *finished = 1 when no tree could split (boosting.rs:1118-1126). */
lgbmhip_status lgbmhip_train_iteration(lgbmhip_ctx* ctx, int32_t* finished);
/* One tree from caller gradients (length num_rows); no score update. */
lgbmhip_status lgbmhip_train_tree(lgbmhip_ctx* ctx, const float* grad, const float* hess);
lgbmhip_status lgbmhip_sync(lgbmhip_ctx* ctx);

int32_t lgbmhip_num_trees(lgbmhip_ctx* ctx);
int32_t lgbmhip_tree_num_leaves(lgbmhip_ctx* ctx, int32_t t);
/* Tree t's arrays; node arrays hold num_leaves-1 entries, leaf arrays num_leaves. */
lgbmhip_status lgbmhip_tree_arrays(lgbmhip_ctx* ctx, int32_t t, int32_t* split_feature, uint32_t* threshold_in_bin,
                                   double* threshold, int8_t* decision_type, int32_t* left_child, int32_t* right_child,
                                   float* split_gain, double* leaf_value, int32_t* leaf_count, double* leaf_weight,
                                   double* internal_value);
/* Leaf index of every training row for the last tree (partition state). */
lgbmhip_status lgbmhip_leaf_of_row(lgbmhip_ctx* ctx, int32_t* out);
lgbmhip_status lgbmhip_copy_scores(lgbmhip_ctx* ctx, double* out);
/* score[i] += leaf_values[leaf of row i in the last tree grown], for every
 * training row (score: host, num_rows values; leaf_values: at least the
 * tree's num_leaves values). */
lgbmhip_status lgbmhip_add_leaf_outputs(lgbmhip_ctx* ctx, const double* leaf_values, int32_t num_values,
                                        double* score);
lgbmhip_status lgbmhip_copy_gradients(lgbmhip_ctx* ctx, float* grad, float* hess);
lgbmhip_status lgbmhip_quant_state(lgbmhip_ctx* ctx, int8_t* gq, int8_t* hq, double* scales);

/* ---- stage entry points (parity checks) ---- */
lgbmhip_status lgbmhip_gradients(lgbmhip_ctx* ctx, const double* scores, float* grad, float* hess);
lgbmhip_status lgbmhip_discretize(lgbmhip_ctx* ctx, const float* grad, const float* hess, int32_t num_calls,
                                  int8_t* gq, int8_t* hq, double* scales);
/* Histogram over rows[0..m) of the uploaded matrix: out is [F][B][2] f64
 * (B = max num_bin), most-frequent bins left at zero (filled later from totals). */
lgbmhip_status lgbmhip_build_histogram(lgbmhip_ctx* ctx, const uint32_t* rows, int64_t m, const float* grad,
                                       const float* hess, double* out, int32_t* out_num_bins);
lgbmhip_status lgbmhip_build_histogram_int(lgbmhip_ctx* ctx, const uint32_t* rows, int64_t m, const int8_t* gq,
                                           const int8_t* hq, int64_t* out, int32_t* out_num_bins);
/* Threshold search on one feature's bin-indexed [B][2] histogram (mfb fixed from the sums first).
 * out_d = {gain, left_output, right_output, lsg, lsh, rsg, rsh}; out_i = {threshold, default_left,
 * left_count, right_count, splittable}; out_cat_bits (8 words, may be NULL) = the bins a categorical
 * split sends left, as a bitset (zero for numerical features). */
lgbmhip_status lgbmhip_find_best_split(lgbmhip_ctx* ctx, int32_t feature, const double* hist, double sum_g,
                                       double sum_h, int32_t num_data, double parent_output, double* out_d,
                                       int32_t* out_i, uint32_t* out_cat_bits);
lgbmhip_status lgbmhip_find_best_split_int(lgbmhip_ctx* ctx, int32_t feature, const int64_t* hist, int64_t int_sum,
                                           double grad_scale, double hess_scale, int32_t num_data,
                                           double parent_output, double* out_d, int32_t* out_i, int64_t* out_gh,
                                           uint32_t* out_cat_bits);
/* Stable partition of rows[0..m) on inner feature `feature`: left rows (in order) then right rows. */
lgbmhip_status lgbmhip_partition(lgbmhip_ctx* ctx, const uint32_t* rows, int64_t m, int32_t feature,
                                 uint32_t threshold, int32_t default_left, uint32_t* out_rows, int64_t* out_left);
/* lgbmhip_partition on device memory that allocates nothing: d_rows[0..m)
 * (device u32 row indices) partitioned stably into d_out_rows (device, m
 * entries: left rows in order, then right rows in order); *out_left (host)
 * receives the left count. No counterpart in the reference; it is for callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc): its tile counts live in a static
 * 256 KiB device buffer, one per GPU, so m is at most 67,106,816 rows
 * (InvalidParameter above). Every call on a GPU shares that buffer: calls that
 * overlap on one GPU (two contexts, or two host threads) race. A row index
 * >= the uploaded rows is InvalidData; numerical features only. */
lgbmhip_status lgbmhip_partition_device_noalloc(lgbmhip_ctx* ctx, const uint32_t* d_rows, int64_t m, int32_t feature,
                                                uint32_t threshold, int32_t default_left, uint32_t* d_out_rows,
                                                int64_t* out_left);
/* Raw scores of the trained model on raw rows (rows x ncol f32): out [rows][K]. */
lgbmhip_status lgbmhip_predict_raw(lgbmhip_ctx* ctx, const float* X, int64_t rows, int32_t ncol, double* out);
/* Raw scores on binned rows (rows x F u8, the mappers of the training data): out [rows][K]. */
lgbmhip_status lgbmhip_predict_binned(lgbmhip_ctx* ctx, const uint8_t* bins, int64_t rows, double* out);
/* Prediction over rows already resident in device memory (exactly one of d_X:
 * rows x ncol f32 raw values, or d_bins: rows x F u8 bins) into a device
 * [rows][K] f64 buffer; returns when the output is written. No counterpart in
 * the reference, whose prediction takes host arrays (predict.rs:231-339); it
 * is the same computation without the host copies, for callers whose rows
 * already live on the GPU. */
lgbmhip_status lgbmhip_predict_device(lgbmhip_ctx* ctx, const float* d_X, const uint8_t* d_bins, int64_t rows,
                                      int32_t ncol, double* d_out);

/* A validation set binned with the training data's mappers (rows x F u8,
 * the layout of lgbmhip_predict_binned): labels (weights, query boundaries
 * and class-major init scores [K][rows] may be NULL). Its scores start from
 * init_score (or zeros), take every tree trained so far, one at a time in
 * model order, and then every later tree on the device as the training
 * scores do. *out_index receives its data index for lgbmhip_eval (1, 2, ...;
 * 0 is the training data). */
lgbmhip_status lgbmhip_add_valid(lgbmhip_ctx* ctx, const uint8_t* bins, int64_t rows, const float* label,
                                 const float* weight, const int32_t* query_boundaries, int32_t num_queries,
                                 const double* init_score, int32_t* out_index);
/* Scores of validation set `index` (1-based), class-major [K][rows]. */
lgbmhip_status lgbmhip_copy_valid_scores(lgbmhip_ctx* ctx, int32_t index, double* out);
/* One metric over the scores of data set `data` (0 training, i >= 1 the
 * i-th validation set), computed on the device. out receives one value, or
 * one per cut-off for NDCG; *out_count (may be NULL) the number written.
 * Unweighted AUC, binary_error, multi_error and NDCG equal the reference's
 * single-thread values exactly; the logloss metrics and weighted AUC sum in
 * a fixed order of their own (within 1e-10 relative). */
lgbmhip_status lgbmhip_eval(lgbmhip_ctx* ctx, int32_t data, const lgbmhip_metric_opts* opts, double* out,
                            int32_t* out_count);

/* PredictKind::LeafIndex (predict.rs:272-279): each raw row's leaf in every
 * tree of the prediction window, out [rows][trees] as f64 (out must hold
 * rows x the window's trees); *out_trees (may be NULL) receives the tree count. */
lgbmhip_status lgbmhip_predict_leaf(lgbmhip_ctx* ctx, const float* X, int64_t rows, int32_t ncol, double* out,
                                    int32_t* out_trees);
/* Gbdt::rollback_one_iter (boosting.rs:1215-1239): the last iteration's trees
 * leave the training and validation scores and the model. */
lgbmhip_status lgbmhip_rollback_one_iter(lgbmhip_ctx* ctx);

/* The prediction window and early stopping of later predict_* calls on this
 * context (Gbdt::predict_window / prediction_early_stop, predict.rs:178-202):
 * iterations start_iteration .. start_iteration + num_iteration (<= 0: to
 * the end); early_stop_freq > 0 stops a row's sum once, at a multiple of
 * freq iterations, its margin (2|raw| or the top-two gap of the classes)
 * exceeds early_stop_margin. Defaults: 0, 0, 0 (every tree, no stopping). */
lgbmhip_status lgbmhip_set_predict_window(lgbmhip_ctx* ctx, int32_t start_iteration, int32_t num_iteration,
                                          int32_t early_stop_freq, double early_stop_margin);

/* Model text (GBDT::SaveModelToString, io.rs:13-88): header, tree_sizes,
 * every tree's block (tree.rs:718-770, numbers as C's %.17g / %g), "end of
 * trees", the split-count feature importances and, when there is one, the
 * parameters block. feature_names (space-separated, one per column),
 * feature_infos (as lgbmhip_bins_feature_infos gives them) and parameters
 * (the host config's `[name: value]` lines, config.rs:1169-1187) may be
 * NULL: a model loaded from text then keeps its own, a trained one writes
 * Column_0 ..., "none" per column and no parameters block. num_iteration
 * <= 0: every tree. *out_len receives the text's length; out (capacity
 * bytes) receives the NUL-terminated text when it fits (out = NULL only
 * sizes it). */
lgbmhip_status lgbmhip_save_model(lgbmhip_ctx* ctx, int32_t num_iteration, const char* feature_names,
                                  const char* feature_infos, const char* parameters, char* out, int64_t capacity,
                                  int64_t* out_len);
/* Trees of a model text (GBDT::LoadModelFromString, io.rs:156-343) into an
 * empty context, for raw prediction (lgbmhip_predict_raw / _device with X);
 * linear trees, forests (average_output) and objectives other than binary,
 * multiclass and lambdarank are Unsupported. */
lgbmhip_status lgbmhip_load_model(lgbmhip_ctx* ctx, const char* text);
/* The parse and checks of lgbmhip_load_model without a context or device
 * (io.rs:156-343, tree.rs:869-1047): field value counts, numbers as
 * parse_f64 / atof_legacy read them (fmt.rs:57-150), child ranges, each node
 * referenced once, categorical indices and bitset boundaries. Out pointers
 * may be NULL; out / out_len receive the text lgbmhip_save_model would write
 * for the parsed model (as lgbmhip_save_model's out / out_len). */
lgbmhip_status lgbmhip_check_model(const char* text, int32_t* out_num_trees, int32_t* out_num_tree_per_iteration,
                                   char* out, int64_t capacity, int64_t* out_len);
/* Dataset::feature_infos (dataset.rs:770-776): per raw column "[min:max]",
 * the categories of a categorical column joined by ':', or "none". */
lgbmhip_status lgbmhip_bins_feature_infos(const lgbmhip_bins* b, char* out, int64_t capacity, int64_t* out_len);

/* Profiling of training iterations: enable = 1 records HIP events around every
 * histogram, slab-reduction, split and partition launch (graphs off), 0 stops,
 * -1 only reads. out12 (may be NULL) receives the totals since enabling:
 * {histogram-kernel ms, histogram launches (one per growth round plus the
 * root's), rows summed into histograms, algorithmic histogram bytes, split ms,
 * partition-kernel ms (the round's plan excluded), iteration ms, iterations, rows moved
 * by partitions, splits computed, splits in the trees, slab-reduction ms}.
 * Each span runs from the previous event on the same stream, so it includes
 * that launch's dispatch gap. No counterpart in the reference (it has no
 * profiler hooks); for bench and rocprof cross-checks. */
lgbmhip_status lgbmhip_profile(lgbmhip_ctx* ctx, int32_t enable, double* out12);

/* The device's LDS-atomic peak: conflict-free ds_add_u32 (each lane of a
 * wave adds to its own word) from three 512-thread blocks per CU, timed with
 * HIP events on the context's stream. out: lane atomics per second. No
 * counterpart in the reference; the denominator of the histogram kernel's
 * LDS-atomic throughput in bench reports. */
lgbmhip_status lgbmhip_lds_atomic_peak(lgbmhip_ctx* ctx, double* lane_atomics_per_sec);

/* Launch shapes of the tree-growth kernels (0 keeps the current value):
 * partition blocks, histogram row chunks, minimum rows per histogram chunk,
 * LDS budget in bytes of one histogram block (sets features per block),
 * growth rounds in a tree's first launch and in each continuation launch,
 * and late_round (-1 keeps the current value, 0 never): from that growth
 * round on, the split search sums the histogram pass's row-chunk units
 * itself and the round has no slab-reduction launch. Defaults: part_grid =
 * min(tiles, 4 x CUs), hist_chunks = 2 x CUs (quantized) or CUs (FP64),
 * 4096 rows (quantized: at most one epoch of the 32-bit histogram bins,
 * 4095 for 16 quantization bins), 32 KiB, 24 and 8 rounds, late_round 12
 * (two more per doubling of rows past 16M). Trees do not depend on any of
 * them.
 * No counterpart in the reference (its thread count is its only knob); for
 * launch-shape sweeps on other row/feature shapes (Booster.set_tuning). */
lgbmhip_status lgbmhip_set_tuning(lgbmhip_ctx* ctx, int32_t part_grid, int32_t hist_chunks, int32_t hist_min_rows,
                                  int32_t hist_lds_budget, int32_t rounds_main, int32_t rounds_cont,
                                  int32_t late_round);
/* The autotuner: before the first iteration, times boosting iterations of
 * the workload itself (trees do not depend on launch shapes) at candidate
 * launch shapes, by coordinate descent over the histogram LDS budget
 * (features per block and the row layout), histogram row chunks, minimum
 * rows per chunk and the partition grid: each candidate runs two warm and
 * trees_per_candidate timed iterations from the starting scores, the
 * training state is reset after each, and the fastest is kept (a change
 * must win by 2%). The choice is cached per
 * process for the workload shape (rows, used features, bins, form,
 * classes, leaves, CUs): a second context of the same shape takes it
 * without timing. out5 (may be NULL): {part_grid, hist_chunks,
 * hist_min_rows, hist_lds_budget (0: default), candidates timed (0: from
 * the cache)}. No counterpart in the reference, whose only shape rule is
 * its row-block size from the thread count (multi_val_bin.rs:337-351).
 * Unsupported with bagging or GOSS. */
lgbmhip_status lgbmhip_autotune(lgbmhip_ctx* ctx, int32_t trees_per_candidate, int32_t* out5);
/* Current launch shapes: {part_grid, hist_chunks, hist_min_rows, hist_lds_budget, features per hist block,
 * hist feature groups, threads per row, rounds_main, rounds_cont, late_round}. */
lgbmhip_status lgbmhip_get_tuning(lgbmhip_ctx* ctx, int32_t* out10);

/* Device-time of the last training iteration's kernels, by kind (ms): histogram, split, partition, other. */
lgbmhip_status lgbmhip_last_timings(lgbmhip_ctx* ctx, double* out4);

#ifdef __cplusplus
}
#endif

#endif /* LGBM_HIP_H_ */
