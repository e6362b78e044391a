"""Benchmark: boosting iterations/sec of the int8-quantized and plain FP64
histogram forms on one MI355X, with histogram GB/s against the HBM roofline.

Workload (default): binary logloss GBDT on 10,500,000 x 28 synthetic rows
(20 N(0,1) + 8 lognormal features, 5% label flips), num_leaves=255,
max_bin=255, min_data_in_leaf=20, learning_rate=0.1. The quantized form
uses 16 int8 gradient levels with stochastic rounding and FP64 leaf renewal.

A step is one boosting iteration (gradients, one 255-leaf tree, score
update) over the full matrix resident in HBM. Warmup iterations run first;
then exactly --steps iterations are timed between device synchronizations.

Beside the timed steps, each form also trains a fresh booster for the full
--rounds (500) with its train loop timed end to end, and reports the
holdout AUC at that last round (|delta AUC| between the forms); the same is
done at each --also-max-bin value (63). The CPU baseline is the oracle's
threaded C++ engine on all rows, at the host's cores (at most 16) and at one
thread, best of --cpu-repeats iterations each.

Prints one JSON line (the driver's contract) and exits 0.

Kernel timing: the timed iterations run as captured HIP graphs with no
instrumentation; --prof-iters further iterations then run with HIP events
around every partition / histogram / slab-reduction / split launch on the
stream the kernels run on, which gives each kernel's average launch duration
(its own launch plus the dispatch gap before it). `roofline` is the dominant
kernel, the partition pass, with the algorithmic bytes of an index partition
(9 per moved row: the split feature's bin, a row index read and written);
`roofline_hist` is the histogram kernel (the metric's hist GB/s). HBM traffic
per launch comes from rocprofv3 FETCH_SIZE / WRITE_SIZE passes over this same
command with --prof-all (every iteration's algorithmic bytes counted), scaled
by the counters' reading on a copy of known size (--copy-calibrate): the
ratios of HBM bytes to algorithmic bytes over the same dispatches are
recorded in profiles/hist_traffic.json, and `traffic` is that ratio times this
run's algorithmic bytes per launch. `iteration` is the whole iteration's
algorithmic bytes over its time.
"""
import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hist_traffic.json")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def holdout_auc(holdout):
    """Quality gate: ROC AUC of the model's raw scores on held-out rows
    (scikit-learn's, independent of the engine)."""
    def q(b):
        from sklearn.metrics import roc_auc_score
        Xv, yv = holdout
        return float(roc_auc_score(yv, b.predict(Xv, raw_score=True)))
    return q


def ndcg_at(scores, rel, qb, k):
    """Mean NDCG@k over queries (gain 2^rel - 1, discount log2(i + 2); a
    query without relevant documents counts 1), score ties broken by document
    order, as a quality report independent of the engine."""
    tot = 0.0
    disc = 1.0 / np.log2(np.arange(2, k + 2))
    for q in range(len(qb) - 1):
        s, r = scores[qb[q]:qb[q + 1]], rel[qb[q]:qb[q + 1]]
        order = np.lexsort((np.arange(len(s)), -s))[:k]
        dcg = float(((2.0 ** r[order] - 1.0) * disc[:len(order)]).sum())
        ideal = np.sort(r)[::-1][:k]
        idcg = float(((2.0 ** ideal - 1.0) * disc[:len(ideal)]).sum())
        tot += dcg / idcg if idcg > 0 else 1.0
    return tot / (len(qb) - 1)


def run_form(lgb, ds, params, warmup, steps, quality, prof_iters, prof_all=False):
    b = lgb.Booster(params, ds)
    if prof_all:
        b.profile(1)   # every iteration's launches counted (counter passes)
    for _ in range(warmup):
        b.update()
    b.update()  # one more untimed iteration (graph already captured)
    t0 = time.perf_counter()
    done = 0
    for _ in range(steps):
        b.update()
        done += 1
    t1 = time.perf_counter()
    its = done / (t1 - t0)
    # kernel-level profile over a few more iterations (HIP events around each launch)
    if not prof_all:
        b.profile(1)
    for _ in range(prof_iters):
        b.update()
    prof = b.profile(0)
    auc = None
    try:
        auc = quality(b)
    except Exception as e:  # a missing metric library is reported, not fatal
        log(f"quality metric unavailable: {e}")
    iters_total = warmup + 1 + steps + prof_iters
    tun = b.get_tuning()
    lds_peak = b.lds_atomic_peak()
    autotuned = b.tuning
    b.free()
    return dict(iters_per_sec=its, ms_per_iter=1000.0 / its, prof=prof, auc=auc, iterations=iters_total,
                hist_fb=int(tun["hist_features_per_block"]), hist_ng=int(tun["hist_groups"]), lds_peak=lds_peak,
                tuning=dict(tun, autotune=autotuned))


def host_cores():
    """Host threads this process may use: its CPU affinity, at most 16 (the
    host share of one GPU on the benchmark machines)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def full_run(lgb, ds, params, rounds, quality):
    """A fresh booster trained for `rounds` iterations, the train loop timed
    end to end (graph capture of the first tree included), then the quality
    metric at that round."""
    b = lgb.Booster(params, ds)
    t0 = time.perf_counter()
    done = 0
    for _ in range(rounds):
        if b.update():
            break
        done += 1
    dt = time.perf_counter() - t0
    q = None
    try:
        q = quality(b)
    except Exception as e:
        log(f"quality metric unavailable: {e}")
    b.free()
    return dict(rounds=done, seconds=dt, iters_per_sec=done / dt, quality=q)


def cpu_baseline(X, y, params, threads, repeats):
    """The repository's CPU engine (the oracle's C++ learner, threaded: its
    trees do not depend on the thread count) on the full workload: one warm
    iteration, then `repeats` timed iterations, the best kept, at `threads`
    host threads and at one thread (the reference's timing method: train loop
    only, best of 3; crates/lgbm-core/benches/train.rs:61-83)."""
    from oracle import pyoracle as O
    t = time.perf_counter()
    ds = O.Dataset(X, y, params)
    log(f"cpu: oracle dataset in {time.perf_counter() - t:.1f}s")
    best = {}
    try:
        for th in (threads, 1):
            O.set_num_threads(th)
            b = O.Booster(ds, params)
            b.update()   # warm: the first iteration includes boost-from-average and the quantizer tables
            ts = []
            for _ in range(repeats):
                t0 = time.perf_counter()
                b.update()
                ts.append(time.perf_counter() - t0)
            best[th] = min(ts)
            log(f"cpu: {th} threads, iterations {['%.3f' % x for x in ts]} s")
            del b
    finally:
        O.set_num_threads(1)
    return dict(value=1.0 / best[threads], unit="iters/sec", cores=threads, kind="port",
                one_thread_iters_per_sec=1.0 / best[1], threaded_speedup=best[1] / best[threads],
                sample=f"the oracle's CPU engine (C++ restatement of the reference learner, quantized form) on all "
                       f"{len(y)} rows x 28 of the same generator, {threads} host threads (one thread beside it): "
                       f"best of {repeats} iterations after one warm iteration")


def copy_calibrate(mib):
    """One device-to-device copy of `mib` MiB (a kernel of known traffic: the
    counters should read mib MiB fetched and mib MiB written)."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(3):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 3
    print(json.dumps({"copy_calibrate_mib": mib, "copies": 3, "ms_per_copy": ms,
                      "copy_gbs_read_plus_write": 2 * mib * (1 << 20) / (ms * 1e-3) / 1e9}), flush=True)


def recorded_traffic(form, rows, max_bin, num_leaves, kernel="hist"):
    """HBM bytes per algorithmic byte of the histogram (or partition)
    launches, from the committed counter passes, when they were taken on
    this exact workload."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    w = t.get("workload", {})
    if (w.get("form"), w.get("rows"), w.get("max_bin"), w.get("num_leaves")) != (form, rows, max_bin, num_leaves):
        return None
    if kernel == "partition":
        return t.get("partition", {}).get("hbm_per_algorithmic_byte")
    return t.get("hbm_per_algorithmic_byte")


# Algorithmic bytes of a partition pass per moved row (the reference's index
# partition: the split feature's bin, a row index read and written); the engine moves each row's
# group-major bins and gradients instead, so every later pass reads its
# leaves contiguously.
PART_ALG_BYTES_PER_ROW = 1 + 4 + 4


def iteration_bytes(p, n, F, B, quant, K=1):
    """Algorithmic bytes of the profiled iterations: histograms, gradients,
    quantization, partitions, score updates and split searches, each at the
    bytes per row or per bin the reference's algorithm must touch."""
    c_out = 8 if quant else 16
    its = max(p["iters"], 1)
    per_iter = (n * K * (8 + 1 + 8)                       # gradients
                + (n * K * (8 + 2 + 16) if quant else 0)  # quantizer (random tables read)
                + n * K * (8 + 8 + 2))                    # score update
    searches = p["splits_computed"] + its * K             # every computed split, and each root
    return p["hist_bytes"] + p["part_rows"] * PART_ALG_BYTES_PER_ROW + per_iter * its + searches * 2 * F * B * c_out


def _line(metric, value, unit, args, ms, dtype, config, **extra):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
           "data": "synthetic", "config": config}
    out.update(extra)
    return out


def _train_forms(lgb, ds, base, args, quality):
    forms = {}
    for form in args.forms.split(","):
        params = dict(base)
        if args.batch_trees is not None:
            params["batch_trees"] = args.batch_trees
        params["pack_u4"] = not args.no_pack_u4
        if form == "quant":
            params.update(use_quantized_grad=True, num_grad_quant_bins=16, stochastic_rounding=True,
                          quant_train_renew_leaf=True)
        r = run_form(lgb, ds, params, args.warmup, args.steps, quality, args.prof_iters)
        log(f"{form}: {r['iters_per_sec']:.2f} it/s ({r['ms_per_iter']:.2f} ms/it), prof {json.dumps(r['prof'])}, "
            f"quality {r['auc']}")
        forms[form] = r
    return forms


def _forms_summary(forms, quality_name):
    return {k: {"iters_per_sec": v["iters_per_sec"], "ms_per_iter": v["ms_per_iter"], quality_name: v["auc"],
                "hist_ms_per_iter": v["prof"]["hist_ms"] / max(v["prof"]["iters"], 1),
                "split_ms_per_iter": v["prof"]["split_ms"] / max(v["prof"]["iters"], 1),
                "slab_ms_per_iter": v["prof"]["slab_ms"] / max(v["prof"]["iters"], 1),
                "part_ms_per_iter": v["prof"]["part_ms"] / max(v["prof"]["iters"], 1)} for k, v in forms.items()}


def workload_multiclass(lgb, datagen, args):
    """10-class softmax, 2M x 100 (8 categorical columns), num_leaves=127:
    one step is one iteration = 10 trees."""
    n, nv = 2_000_000, 100_000
    t = time.perf_counter()
    X, y = datagen.multiclass_2m(n + nv)
    log(f"data {X.shape} in {time.perf_counter() - t:.1f}s")
    Xv, yv = X[n:].copy(), y[n:].copy()
    base = dict(objective="multiclass", num_class=10, num_leaves=127, max_bin=255, min_data_in_leaf=20,
                learning_rate=0.1, categorical_feature=datagen.MULTICLASS_CATEGORICAL)
    ds = lgb.Dataset(X[:n], label=y[:n], params=base)
    del X

    def accuracy(b):
        return float((b.predict(Xv).argmax(1) == yv).mean())

    forms = _train_forms(lgb, ds, base, args, accuracy)
    main_form = "quant" if "quant" in forms else list(forms)[0]
    m = forms[main_form]
    return _line("boosting iters/sec (10-class softmax, 10 trees per iteration)", m["iters_per_sec"], "iters/sec",
                 args, m["ms_per_iter"], "int64" if main_form == "quant" else "f64",
                 {"workload": f"multiclass softmax GBDT, {n} x 100 (80 N(0,1), 12 low-cardinality, 8 categorical "
                              f"Zipf(1.2) columns), num_class=10, num_leaves=127, max_bin=255, lr=0.1",
                  "rows": n, "features": 100, "form": main_form, "parallelism": "1 GPU"},
                 forms=_forms_summary(forms, "holdout_accuracy"))


def workload_rank(lgb, datagen, args):
    """Lambdarank NDCG, ~2M documents in ~20k queries x 136, num_leaves=255."""
    t = time.perf_counter()
    X, rel, qb = datagen.rank_2m()
    log(f"data {X.shape}, {len(qb) - 1} queries in {time.perf_counter() - t:.1f}s")
    base = dict(objective="lambdarank", num_leaves=255, max_bin=255, min_data_in_leaf=20, learning_rate=0.1)
    ds = lgb.Dataset(X, label=rel, group=np.diff(qb), params=base)
    nq = 2000
    Xs, rs, qs = X[:qb[nq]].copy(), rel[:qb[nq]].copy(), qb[:nq + 1].copy()
    del X

    def ndcg(b):
        return ndcg_at(b.predict(Xs, raw_score=True), rs, qs, 10)

    forms = _train_forms(lgb, ds, base, args, ndcg)
    main_form = "quant" if "quant" in forms else list(forms)[0]
    m = forms[main_form]
    return _line("boosting iters/sec (lambdarank NDCG)", m["iters_per_sec"], "iters/sec", args, m["ms_per_iter"],
                 "int64" if main_form == "quant" else "f64",
                 {"workload": f"lambdarank GBDT, {int(qb[-1])} documents in {len(qb) - 1} queries x 136, "
                              f"num_leaves=255, max_bin=255, lr=0.1",
                  "rows": int(qb[-1]), "features": 136, "form": main_form, "parallelism": "1 GPU"},
                 forms=_forms_summary(forms, "train_ndcg10_first_2000_queries"))


def workload_predict(lgb, datagen, args):
    """Prediction of a 1000-tree, 255-leaf lambdarank model on 1,000,000 rows
    x 136 already resident in HBM: raw f32 values and u8 bins. One step is
    one pass over all rows."""
    import torch
    trees, rows = 1000, 1_000_000
    t = time.perf_counter()
    X, rel, qb = datagen.rank_2m()
    nq = int(np.searchsorted(qb, 200_000))
    base = dict(objective="lambdarank", num_leaves=255, max_bin=255, min_data_in_leaf=20, learning_rate=0.1,
                use_quantized_grad=True, num_grad_quant_bins=16, quant_train_renew_leaf=True)
    dtrain = lgb.Dataset(X[:qb[nq]], label=rel[:qb[nq]], group=np.diff(qb[:nq + 1]), params=base)
    b = lgb.Booster(base, dtrain)
    for _ in range(trees):
        b.update()
    leaves = [b.tree(i).num_leaves for i in range(trees)]
    log(f"model: {trees} trees, mean {np.mean(leaves):.1f} leaves, trained in {time.perf_counter() - t:.1f}s")
    Xp = np.ascontiguousarray(X[:rows])
    bins = lgb.Dataset(Xp, reference=dtrain).bins
    F = bins.shape[1]
    res = {}
    for form, host, elem in (("raw", Xp, 4), ("binned", bins, 1)):
        dx = torch.from_numpy(host).cuda()
        out = torch.empty(rows, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        for _ in range(args.warmup):
            b.predict_device(dx.data_ptr(), rows, host.shape[1], out.data_ptr(), binned=form == "binned")
        t0 = time.perf_counter()
        for _ in range(args.steps):
            b.predict_device(dx.data_ptr(), rows, host.shape[1], out.data_ptr(), binned=form == "binned")
        dt = (time.perf_counter() - t0) / args.steps
        ref = np.ravel(b.predict(Xp[:20000], raw_score=True))
        got = out[:20000].cpu().numpy()
        byts = rows * host.shape[1] * elem + rows * 8
        res[form] = dict(rows_per_sec=rows / dt, ms_per_pass=dt * 1e3, gbs=byts / dt / 1e9,
                         max_abs_diff_vs_host_predict=float(np.max(np.abs(got - ref))))
        log(f"predict {form}: {rows / dt / 1e6:.2f} M rows/s, {dt * 1e3:.3f} ms/pass, {byts / dt / 1e9:.1f} GB/s")
        del dx, out
    m = res["binned"]
    return _line("prediction rows/sec (1000 trees x 255 leaves, u8 bins; raw f32 form beside it)",
                 m["rows_per_sec"], "rows/sec", args, m["ms_per_pass"], "u8",
                 {"workload": f"predict {rows} x 136 rows (lambdarank feature generator) with a {trees}-tree, "
                              f"255-leaf model, inputs resident in HBM", "rows": rows, "features": 136,
                  "trees": trees, "inner_features": F, "parallelism": "1 GPU"},
                 forms=res,
                 roofline=dict(bound="hbm", achieved=m["gbs"], peak=HBM_PEAK_GBS, unit="GB/s",
                               frac=m["gbs"] / HBM_PEAK_GBS, traffic=None, kernel="k_predict<binned>"))


def workload_eval(lgb, datagen, args):
    """Boosting with evaluation on the GPU: the headline quantized binary
    model on --rows x 28 with a 1,000,000-row validation set attached, whose
    scores every tree updates on the device. One step is one iteration plus
    the validation AUC (what early stopping reads each iteration). Also
    reports AUC and logloss over the training scores, and the oracle's AUC
    (one CPU thread, metric.rs:701-729) over the same training scores."""
    n, nv = args.rows, 1_000_000
    t = time.perf_counter()
    X, y = datagen.binary_10p5m(n)
    Xv, yv = datagen.holdout("binary_10p5m", nv)
    log(f"data {X.shape} + valid {Xv.shape} in {time.perf_counter() - t:.1f}s")
    base = dict(objective="binary", num_leaves=args.num_leaves, max_bin=args.max_bin, min_data_in_leaf=20,
                learning_rate=0.1, use_quantized_grad=True, num_grad_quant_bins=16, stochastic_rounding=True,
                quant_train_renew_leaf=True, metric="auc")
    ds = lgb.Dataset(X, label=y, params=base)
    dv = lgb.Dataset(Xv, label=yv, reference=ds, params=base)
    del X
    b = lgb.Booster(base, ds)
    b.add_valid(dv, "valid")
    for _ in range(args.warmup):
        b.update()
        b.eval_valid()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.update()
        res = b.eval_valid()
    dt = (time.perf_counter() - t0) / args.steps
    log(f"iteration + valid AUC: {1 / dt:.2f} it/s ({dt * 1e3:.2f} ms), valid auc {res[0].metric_value:.6f}")
    timings = {}
    for m in ("auc", "binary_logloss"):
        b._metrics = [m]
        b.eval_train()
        reps = 20
        t0 = time.perf_counter()
        for _ in range(reps):
            v = b.eval_train()[0].metric_value
        timings[m] = dict(ms=(time.perf_counter() - t0) / reps * 1e3, value=v)
        log(f"train {m} over {n} rows: {timings[m]['ms']:.2f} ms ({v:.9f})")
    out = dict(iters_with_valid_auc_per_sec=1 / dt, ms_per_iter=dt * 1e3, valid_auc=res[0].metric_value,
               train_metric_ms=timings)
    if not args.no_cpu:
        from oracle import pyoracle as O
        O.build()
        ts = b.train_scores()
        t0 = time.perf_counter()
        ref = O.metric_auc(ts, y)
        cpu_s = time.perf_counter() - t0
        out["cpu_auc_s"] = cpu_s
        out["cpu_auc_equal"] = bool(ref == timings["auc"]["value"])
        log(f"oracle AUC over {n} rows: {cpu_s:.2f} s, equal: {out['cpu_auc_equal']}")
    auc_ms = timings["auc"]["ms"]
    byts = n * (8 + 4)   # one read of the scores and labels
    return _line("boosting iters/sec with a 1M-row validation set and its AUC every iteration", 1 / dt, "iters/sec",
                 args, dt * 1e3, "int64",
                 {"workload": f"binary logloss GBDT, {n} x 28, num_leaves={args.num_leaves}, max_bin={args.max_bin}, "
                              f"quantized (16 levels), validation set {nv} rows, metric=auc",
                  "rows": n, "valid_rows": nv, "parallelism": "1 GPU"},
                 eval=out,
                 roofline=dict(bound="hbm", achieved=byts / auc_ms / 1e6, peak=HBM_PEAK_GBS, unit="GB/s",
                               frac=byts / auc_ms / 1e6 / HBM_PEAK_GBS, traffic=None,
                               kernel="AUC over the training scores (radix sort of 64-bit keys and runs)"))


def workload_binning(lgb, datagen, args):
    """Dataset construction for --rows x 28 f32 rows: the host's binning
    (mappers from a 200,000-row sample, then every value binned on host
    threads) against the mappers on the host and the binning pass on the
    device, from rows in host memory (with their PCIe copy) and from rows
    already in HBM. One step is one binning pass over all rows."""
    import ctypes as C
    import torch
    n = args.rows
    X, y = datagen.binary_10p5m(n)
    p = dict(objective="binary", max_bin=args.max_bin, num_leaves=args.num_leaves)
    t0 = time.perf_counter()
    host = lgb.Dataset(X, label=y, params=p)
    host_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    dev = lgb.Dataset(X, label=y, params=dict(p, device_binning=True))
    mapper_s = time.perf_counter() - t0
    b = lgb.Booster(p, dev)   # uploads through the device binning pass
    lib, check = lgb._capi.lib(), lgb._capi.check
    t0 = time.perf_counter()
    for _ in range(3):
        check(lib.lgbmhip_upload_raw(b._ctx, dev._handle, lgb.ptr(X), n, X.shape[1], 0), b._ctx)
    from_host_s = (time.perf_counter() - t0) / 3
    dx = torch.from_numpy(X).cuda()
    out = torch.empty((n, dev.num_features), dtype=torch.uint8, device="cuda")
    for _ in range(args.warmup):
        b.bin_rows_device(dx.data_ptr(), n, X.shape[1], out.data_ptr())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.bin_rows_device(dx.data_ptr(), n, X.shape[1], out.data_ptr())
    dt = (time.perf_counter() - t0) / args.steps
    same = bool(np.array_equal(out[:200000].cpu().numpy(), host.bins[:200000]))
    byts = n * X.shape[1] * 4 + n * dev.num_features
    log(f"host binning {host_s:.2f}s; mappers {mapper_s:.2f}s + device pass from host rows {from_host_s * 1e3:.1f} ms; "
        f"device pass on HBM rows {dt * 1e3:.3f} ms ({byts / dt / 1e9:.0f} GB/s), equal: {same}")
    return _line("rows binned per second (device value_to_bin pass, rows resident in HBM)", n / dt, "rows/sec", args,
                 dt * 1e3, "f32->u8",
                 {"workload": f"binning {n} x 28 f32 rows (20 N(0,1) + 8 lognormal) with max_bin={args.max_bin} "
                              f"mappers", "rows": n, "features": 28, "parallelism": "1 GPU"},
                 binning=dict(host_total_s=host_s, host_mappers_s=mapper_s, device_from_host_rows_ms=from_host_s * 1e3,
                              device_from_hbm_rows_ms=dt * 1e3, bins_equal_host=same),
                 roofline=dict(bound="hbm", achieved=byts / dt / 1e9, peak=HBM_PEAK_GBS, unit="GB/s",
                               frac=byts / dt / 1e9 / HBM_PEAK_GBS, traffic=None, kernel="k_bin_rows"))


def workload_bagging(lgb, datagen, args, goss=False):
    """The headline binary workload with row bagging (bagging_fraction=0.5,
    bagging_freq=1: a new bag every iteration, trees grown on its rows,
    out-of-bag rows scored by the tree walk), or with GOSS (top_rate 0.2,
    other_rate 0.1, 4096 sampling blocks; warm-up iterations are part of the
    untimed warmup)."""
    t = time.perf_counter()
    X, y = datagen.binary_10p5m(args.rows)
    Xv, yv = datagen.holdout("binary_10p5m", 200_000)
    log(f"data {X.shape} in {time.perf_counter() - t:.1f}s")
    base = dict(objective="binary", num_leaves=args.num_leaves, max_bin=args.max_bin, min_data_in_leaf=20,
                learning_rate=0.1, bagging_fraction=0.5, bagging_freq=1, bagging_seed=3)
    if goss:
        base.update(bagging_fraction=1.0, bagging_freq=0, data_sample_strategy="goss", top_rate=0.2, other_rate=0.1,
                    num_threads=4096)
        args.warmup = max(args.warmup, 10)
    ds = lgb.Dataset(X, label=y, params=base)
    del X
    forms = _train_forms(lgb, ds, base, args, holdout_auc((Xv, yv)))
    main_form = "quant" if "quant" in forms else list(forms)[0]
    m = forms[main_form]
    what = "GOSS (top_rate 0.2, other_rate 0.1)" if goss else "row bagging (fraction 0.5, a new bag each iteration)"
    cfg = "GOSS top_rate=0.2 other_rate=0.1" if goss else "bagging_fraction=0.5, bagging_freq=1"
    return _line(f"boosting iters/sec with {what}", m["iters_per_sec"],
                 "iters/sec", args, m["ms_per_iter"], "int64" if main_form == "quant" else "f64",
                 {"workload": f"binary logloss GBDT, {args.rows} x 28, num_leaves={args.num_leaves}, "
                              f"max_bin={args.max_bin}, {cfg}",
                  "rows": args.rows, "features": 28, "form": main_form, "parallelism": "1 GPU"},
                 forms=_forms_summary(forms, "holdout_auc"))


WORKLOADS = {"binary_10p5m": None, "multiclass_2m": workload_multiclass, "rank_2m": workload_rank,
             "predict_1m": workload_predict, "eval_10p5m": workload_eval, "binning_10p5m": workload_binning,
             "bagging_10p5m": workload_bagging,
             "goss_10p5m": lambda lgb, datagen, args: workload_bagging(lgb, datagen, args, goss=True)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--rows", type=int, default=10_500_000)
    ap.add_argument("--max-bin", type=int, default=255)
    ap.add_argument("--num-leaves", type=int, default=255)
    ap.add_argument("--prof-iters", type=int, default=3)
    ap.add_argument("--cpu-repeats", type=int, default=3, help="timed CPU iterations per thread count (best kept)")
    ap.add_argument("--rounds", type=int, default=500,
                    help="full training run per form (train loop timed, holdout AUC at its last round); 0 skips")
    ap.add_argument("--also-max-bin", default="63", help="further max_bin values measured the same way")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="quant,plain")
    ap.add_argument("--no-graph", action="store_true", help="stream launches instead of graphs (counter passes)")
    ap.add_argument("--no-autotune", action="store_true", help="default launch shapes instead of the autotuner's")
    ap.add_argument("--no-pack-u4", action="store_true",
                    help="features of at most 16 bins stored as bytes instead of 4-bit fields")
    ap.add_argument("--batch-trees", type=int, default=None,
                    help="multiclass: class trees per shared growth launch (0: one tree per launch)")
    ap.add_argument("--prof-all", action="store_true",
                    help="count every iteration's histogram bytes, warm-up included (counter passes)")
    ap.add_argument("--copy-calibrate", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (counter scaling) and exit")
    ap.add_argument("--workload", default="binary_10p5m", choices=list(WORKLOADS),
                    help="binary_10p5m (the headline line); the others print their own line")
    args = ap.parse_args()
    if args.copy_calibrate:
        copy_calibrate(args.copy_calibrate)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported (the multi-GPU row-sharded "
              "path is a design, not a build)", file=sys.stderr)
        sys.exit(2)

    lgb = importlib.import_module("lightgbm-rust_amd")
    datagen = importlib.import_module("lightgbm-rust_amd.datagen")
    if args.workload != "binary_10p5m":
        print(json.dumps(WORKLOADS[args.workload](lgb, datagen, args)), flush=True)
        return
    t = time.perf_counter()
    X, y = datagen.binary_10p5m(args.rows)
    Xv, yv = datagen.holdout("binary_10p5m", 200_000)
    log(f"data {X.shape} in {time.perf_counter() - t:.1f}s")
    base = dict(objective="binary", num_leaves=args.num_leaves, max_bin=args.max_bin, min_data_in_leaf=20,
                learning_rate=0.1, device_type="gpu", use_graph=not args.no_graph, autotune=not args.no_autotune)
    t = time.perf_counter()
    ds = lgb.Dataset(X, label=y, params=base)
    log(f"host binning in {time.perf_counter() - t:.1f}s, {ds.num_features} features")
    forms = {}
    datasets = {args.max_bin: ds}
    extra_bins = [int(v) for v in args.also_max_bin.split(",") if v] if args.also_max_bin else []
    for mb in extra_bins:
        if mb not in datasets:
            t = time.perf_counter()
            datasets[mb] = lgb.Dataset(X, label=y, params=dict(base, max_bin=mb))
            log(f"host binning max_bin={mb} in {time.perf_counter() - t:.1f}s")
    runs = [(f, args.max_bin) for f in args.forms.split(",")] + \
           [(f, mb) for mb in extra_bins for f in args.forms.split(",")]
    for form, mb in runs:
        params = dict(base, max_bin=mb)
        if form == "quant":
            params.update(use_quantized_grad=True, num_grad_quant_bins=16, stochastic_rounding=True,
                          quant_train_renew_leaf=True)
        key = form if mb == args.max_bin else f"{form}_max_bin_{mb}"
        r = run_form(lgb, datasets[mb], params, args.warmup, args.steps, holdout_auc((Xv, yv)), args.prof_iters,
                     args.prof_all)
        p = r["prof"]
        hist_gbs = p["hist_bytes"] / (p["hist_ms"] * 1e-3) / 1e9 if p["hist_ms"] > 0 else None
        r["hist_gbs"] = hist_gbs
        r["hist_avg_launch_us"] = 1000.0 * p["hist_ms"] / max(p["hist_launches"], 1)
        r["hist_bytes_per_launch"] = p["hist_bytes"] / max(p["hist_launches"], 1)
        r["max_bin"] = mb
        log(f"{key}: {r['iters_per_sec']:.2f} it/s ({r['ms_per_iter']:.2f} ms/it), hist {hist_gbs} GB/s, "
            f"prof {json.dumps(p)}, auc {r['auc']}")
        if args.rounds > 0:
            fr = full_run(lgb, datasets[mb], params, args.rounds, holdout_auc((Xv, yv)))
            r["full"] = fr
            log(f"{key}: {fr['rounds']} rounds in {fr['seconds']:.2f}s ({fr['iters_per_sec']:.2f} it/s), "
                f"holdout auc {fr['quality']}")
        forms[key] = r

    main_form = "quant" if "quant" in forms else list(forms)[0]
    m = forms[main_form]
    p = m["prof"]
    ratio = recorded_traffic(main_form, args.rows, args.max_bin, args.num_leaves)
    traffic = ratio * m["hist_bytes_per_launch"] if ratio else None
    roof_hist = dict(bound="hbm", achieved=m["hist_gbs"], peak=HBM_PEAK_GBS, unit="GB/s",
                     frac=(m["hist_gbs"] / HBM_PEAK_GBS) if m["hist_gbs"] else None,
                     traffic=traffic, traffic_per_algorithmic_byte=ratio,
                     kernel="k_hist_int" if main_form == "quant" else "k_hist_f64",
                     bytes_per_launch=m["hist_bytes_per_launch"], avg_launch_us=m["hist_avg_launch_us"])
    # the dominant kernel by time: the partition pass (one launch per growth
    # round; every launch but the roots' histograms is a round's)
    part_launches = max(p["hist_launches"] - p["iters"], 1)
    part_bytes = p["part_rows"] * PART_ALG_BYTES_PER_ROW / part_launches
    part_us = 1000.0 * p["part_ms"] / part_launches
    pratio = recorded_traffic(main_form, args.rows, args.max_bin, args.num_leaves, "partition")
    moved = 2 * (m["hist_fb"] * m["hist_ng"] + (2 if main_form == "quant" else 8))   # bins and gradients, read and written
    roof = dict(bound="hbm", achieved=part_bytes / part_us / 1e3 if part_us > 0 else None, peak=HBM_PEAK_GBS,
                unit="GB/s", frac=(part_bytes / part_us / 1e3 / HBM_PEAK_GBS) if part_us > 0 else None,
                traffic=pratio * part_bytes if pratio else None, traffic_per_algorithmic_byte=pratio,
                kernel="k_grow_part", bytes_per_launch=part_bytes, avg_launch_us=part_us,
                moved_bytes_per_row=moved,
                moved_gbs=p["part_rows"] * moved / part_launches / part_us / 1e3 if part_us > 0 else None)
    it_bytes = iteration_bytes(p, args.rows, ds.num_features, args.max_bin, main_form == "quant")
    it_gbs = it_bytes / max(p["iters"], 1) / (m["ms_per_iter"] * 1e-3) / 1e9
    # LDS-atomic throughput of the histogram kernel against the device's
    # conflict-free ds_add_u32 rate: one update per (row, feature) is the
    # algorithmic count; the kernel issues one per (row, feature slot of its
    # groups), the most frequent bin's and padding slots' adds included
    hist_s = p["hist_ms"] * 1e-3
    lds = None
    if hist_s > 0:
        alg = p["hist_rows"] * ds.num_features / hist_s
        lds = dict(achieved=alg, issued=p["hist_rows"] * m["hist_fb"] * m["hist_ng"] / hist_s, peak=m["lds_peak"],
                   unit="lane atomics/s", frac=alg / m["lds_peak"], kernel=roof_hist["kernel"])
    cpu = None
    if not args.no_cpu:
        t = time.perf_counter()
        cparams = {k: v for k, v in base.items() if k not in ("device_type", "use_graph", "autotune")}
        if main_form == "quant":
            cparams.update(use_quantized_grad=True, num_grad_quant_bins=16, stochastic_rounding=True,
                           quant_train_renew_leaf=True)
        try:
            cpu = cpu_baseline(X, y, cparams, host_cores(), args.cpu_repeats)
        except Exception as e:
            log(f"cpu baseline failed: {e}")
        log(f"cpu baseline in {time.perf_counter() - t:.1f}s")
    del X
    out = {
        "metric": "boosting iters/sec (int8-quantized histogram form; FP64 form and hist GB/s beside it)",
        "value": m["iters_per_sec"],
        "unit": "iters/sec",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["ms_per_iter"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64" if main_form == "quant" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"binary logloss GBDT, {args.rows} x 28 (20 N(0,1) + 8 lognormal, 5% label flips), "
                        f"num_leaves={args.num_leaves}, max_bin={args.max_bin}, min_data_in_leaf=20, lr=0.1",
            "rows": args.rows, "features": 28, "num_leaves": args.num_leaves, "max_bin": args.max_bin,
            "form": main_form, "parallelism": "1 GPU",
        },
        "forms": {k: dict(iters_per_sec=v["iters_per_sec"], ms_per_iter=v["ms_per_iter"], hist_gbs=v["hist_gbs"],
                          hist_avg_launch_us=v["hist_avg_launch_us"], auc_holdout=v["auc"],
                          iterations_trained=v["iterations"], max_bin=v["max_bin"], launch_shapes=v["tuning"],
                          full_run=(dict(rounds=v["full"]["rounds"], seconds=v["full"]["seconds"],
                                         iters_per_sec=v["full"]["iters_per_sec"],
                                         auc_holdout=v["full"]["quality"]) if "full" in v else None),
                          split_ms_per_iter=v["prof"]["split_ms"] / max(v["prof"]["iters"], 1),
                          slab_ms_per_iter=v["prof"]["slab_ms"] / max(v["prof"]["iters"], 1),
                          part_ms_per_iter=v["prof"]["part_ms"] / max(v["prof"]["iters"], 1),
                          hist_ms_per_iter=v["prof"]["hist_ms"] / max(v["prof"]["iters"], 1))
                  for k, v in forms.items()},
        "roofline": roof,
        "roofline_hist": roof_hist,
        "lds_atomics": lds,
        "iteration": dict(algorithmic_bytes_per_iter=it_bytes / max(p["iters"], 1), gbs=it_gbs,
                          frac=it_gbs / HBM_PEAK_GBS,
                          note="algorithmic bytes of the profiled iterations over the timed ms per iteration"),
        "cpu_baseline": cpu,
    }
    # quantized against plain at each max_bin: the timed steps, the full runs
    # (iters/sec over all rounds) and |delta AUC| on the holdout at the last round
    cmp = {}
    for mb in [args.max_bin] + extra_bins:
        q, pl = ("quant", "plain") if mb == args.max_bin else (f"quant_max_bin_{mb}", f"plain_max_bin_{mb}")
        if q in forms and pl in forms:
            c = dict(steps_speedup=forms[q]["iters_per_sec"] / forms[pl]["iters_per_sec"])
            if "full" in forms[q] and "full" in forms[pl]:
                fq, fp = forms[q]["full"], forms[pl]["full"]
                c["rounds"] = fq["rounds"]
                c["full_run_speedup"] = fq["iters_per_sec"] / fp["iters_per_sec"]
                if fq["quality"] is not None and fp["quality"] is not None:
                    c["abs_delta_auc_at_last_round"] = abs(fq["quality"] - fp["quality"])
            cmp[f"max_bin_{mb}"] = c
    out["quant_vs_plain"] = cmp
    if "quant" in forms and "plain" in forms:
        out["quant_vs_plain_speedup"] = forms["quant"]["iters_per_sec"] / forms["plain"]["iters_per_sec"]
        if forms["quant"]["auc"] is not None and forms["plain"]["auc"] is not None:
            out["delta_auc"] = forms["quant"]["auc"] - forms["plain"]["auc"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
