"""Benchmark: 50-tree random-forest fit at 581,012 x 54 and batched predict on
one MI355X, with the split-scan / reshuffle roofline and CPU baselines.

    python bench.py [--gpus 1] [--steps 10] [--warmup 3]

A step is one complete fit of 50 trees (device transpose + presort, bootstrap
compaction, every level of every tree, export of the canonical trees into
host memory) on the covtype-shaped synthetic set with its inputs already
resident in HBM; copying the last step's trees into numpy arrays for the
checks is not timed.  The K
timed steps run without per-kernel events (value, ms_per_step); the same K
steps are then repeated with HIP events around every kernel group for the
per-kernel times and the roofline (ms_per_step_with_kernel_events).  Prints
one JSON line.  Only one GPU is used; --gpus above 1 exits with a message.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "fit s/50 trees @581k×54; predict rows/s + p99 latency; HBM GB/s fraction"
HBM_PEAK_GBS = 8000.0
N_TREES = 50
MAX_FEATURES = 7
SEED = 20240601


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def fit_once(ctx, lib, Xd, yd, N, F, K, mode="batched", n_trees=N_TREES, export=True, m=MAX_FEATURES, seed=SEED):
    """One complete fit from the resident inputs to a forest whose canonical
    trees are in host memory; export=True also copies them into numpy arrays."""
    ds_h = lib.load()
    import ctypes

    h = ctypes.c_void_p()
    lib.check(ds_h.cta_dataset_create_device(ctx.h, ctypes.c_void_p(Xd.data_ptr()), ctypes.c_void_p(yd.data_ptr()),
                                             N, F, K, 0, ctypes.byref(h)))
    ds = lib.Dataset.__new__(lib.Dataset)
    ds.ctx, ds.lib, ds.h, ds.N, ds.F, ds.K = ctx, ds_h, h, N, F, K
    o = lib.default_opts()
    o.seed = seed
    o.max_features = m
    o.mode = lib.MODES[mode]
    o.bootstrap = 1  # in-bag sets drawn on the device from (seed, tree)
    forest = lib.Forest.fit(ctx, ds, n_trees, o, None)
    arrays = forest.export() if export else None
    ds.close()
    return forest, arrays


def kernel_table(stats, steps):
    """Per-kernel ms per fit, achieved algorithmic GB/s and fraction of peak
    from the context's HIP-event statistics."""
    out = {}
    for k, v in sorted(stats.items(), key=lambda kv: -kv[1]["ms"]):
        if k.startswith("host_") or v["ms"] <= 0:
            continue
        e = {"ms_per_fit": round(v["ms"] / steps, 3), "launches_per_fit": round(v["launches"] / steps, 1)}
        if v["alg_bytes"]:
            gbs = v["alg_bytes"] / (v["ms"] * 1e-3) / 1e9
            e.update({"achieved_gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)})
        out[k] = e
    return out


# The other fit workloads: full-size synthetic sets of the generators in
# cudatree_amd.datasets, timed as fits from resident inputs.
WORKLOADS = [
    ("gaussian_clusters_200000x256", "200,000 x 256 float32, 20 classes of 3 Gaussian clusters (centroids N(0, 2^2), "
     "sigma 1); 200 trees, bootstrap, max_features=16, full depth", "gaussian_clusters", {}, 200, 16),
    ("zipf_quantized_1000000x64", "1,000,000 x 64 float32 quantised to 256 levels, 100 classes with Zipf(1.1) "
     "frequencies; 100 trees, bootstrap, max_features=8, full depth", "zipf_quantized", {}, 100, 8),
    ("wide_sparse_20000x4096", "20,000 x 4,096 float32, 5 classes, 64 informative features (0.5 sigma shift); "
     "100 trees, bootstrap, max_features=64, full depth", "wide_sparse_signal", {}, 100, 64),
]


def workload_bench(lib, ctx, args, torch, dev):
    """s per fit and per-kernel fractions of the other workloads; the
    autotuner's sweep on the 200,000 x 256 set."""
    from cudatree_amd import autotune, datasets

    out = {}
    for name, desc, gen, kw, T, m in WORKLOADS:
        X, y = getattr(datasets, gen)(**kw)
        N, F = X.shape
        K = int(y.max()) + 1
        Xd = torch.from_numpy(X).to(dev)
        yd = torch.from_numpy(y.astype(np.uint32).view(np.int32)).to(dev)
        torch.cuda.synchronize()
        forest, arrays = fit_once(ctx, lib, Xd, yd, N, F, K, n_trees=T, m=m, seed=SEED + 7)
        nodes = len(arrays["feature"])
        del forest
        steps = args.workload_steps
        ctx.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            forest, _ = fit_once(ctx, lib, Xd, yd, N, F, K, n_trees=T, m=m, seed=SEED + 7, export=False)
            del forest
        ctx.synchronize()
        dt = (time.perf_counter() - t) / steps
        ctx.reset_stats()
        ctx.set_profiling(True)
        forest, _ = fit_once(ctx, lib, Xd, yd, N, F, K, n_trees=T, m=m, seed=SEED + 7, export=False)
        ctx.set_profiling(False)
        del forest
        e = {"workload": desc, "s_per_fit": round(dt, 4), "steps": steps, "n_trees": T, "nodes": nodes,
             "kernels": kernel_table(ctx.kernel_stats(), 1)}
        ctx.reset_stats()
        log("workload %s: %.4f s per fit of %d trees (%d nodes)" % (name, dt, T, nodes))
        if name.startswith("gaussian") and args.tune_probe_trees > 0:
            t = time.perf_counter()
            best, table = autotune.tune(X, y, K, m, probe_trees=args.tune_probe_trees, repeats=1, seed=SEED + 7,
                                        use_cache=False)
            e["autotune"] = {"probe": "%d of %d trees on the full set, one timing per setting" %
                                      (args.tune_probe_trees, T),
                             "best": best, "sweep_s": round(time.perf_counter() - t, 1),
                             "table": [{"setting": st, "s": round(sec, 4)} for st, sec in table]}
            log("autotune %s: best %s (%.4f s) of %d settings" % (name, best, table[0][1], len(table)))
        out[name] = e
        del Xd, yd
        ctx.release_workspace()
        torch.cuda.empty_cache()
    return out


def serving_bench(lib, X, y, Xt, args):
    """Resident forest of the tabular type (serving_trees trees, depth <= 24)
    behind the serving loop.  A sweep of Poisson streams of predict requests
    (log-uniform 1..65,536 rows) mixed with small fit jobs (1,000-200,000
    rows, 10-50 trees), each replayed with coalesced launches (short and long
    request lanes) and with one launch per request, up to rates the
    per-request launches cannot keep up with; then a stream of small requests
    (log-uniform 1..64 rows), where launches rather than traversal set the
    rate."""
    import threading

    ctx = lib.default_context(0)
    ds = lib.Dataset(ctx, X, y, int(y.max()) + 1)
    o = lib.default_opts()
    o.seed = SEED + 1
    o.max_features = MAX_FEATURES
    o.max_depth = 24
    o.bootstrap = 1
    t = time.perf_counter()
    forest = lib.Forest.fit(ctx, ds, args.serving_trees, o, None)
    t_fit = time.perf_counter() - t
    ds.close()
    ctx.release_workspace()
    log("serving forest: %d trees, %d nodes, fit %.2fs" % (forest.n_trees, forest.n_nodes, t_fit))
    dur = args.serving_seconds
    K = int(y.max()) + 1

    def stream(qps, largest, with_fits, coalesce, seed):
        rng = np.random.default_rng(seed)
        n_req = int(qps * dur)
        arrivals = np.cumsum(rng.exponential(1.0 / qps, size=n_req))
        sizes = np.minimum(np.exp(rng.uniform(0, np.log(largest), size=n_req)).astype(int), largest)
        starts = rng.integers(0, len(Xt) - largest, size=n_req)
        fits = []
        if with_fits:
            for k in range(int(dur // 2)):
                n = int(np.exp(rng.uniform(np.log(1000), np.log(200000))))
                fits.append((2.0 * k + 1.0, n, int(rng.integers(10, 51)), int(rng.integers(0, len(X) - n))))
        srv = lib.Server(coalesce=coalesce, fit_workers=1)
        fid = srv.add_forest(forest)
        tickets, fit_tickets = [], []
        t0 = time.perf_counter()

        def producer():
            fi = 0
            for i in range(n_req):
                while fi < len(fits) and fits[fi][0] <= arrivals[i]:
                    _, n, T, a = fits[fi]
                    fo = lib.default_opts()
                    fo.seed = 5 + fi
                    fo.max_features = MAX_FEATURES
                    fo.bootstrap = 1
                    fit_tickets.append(srv.submit_fit(X[a:a + n], y[a:a + n], K, T, fo))
                    fi += 1
                delay = arrivals[i] - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
                tickets.append(srv.submit_predict(fid, Xt[starts[i]:starts[i] + sizes[i]]))

        th = threading.Thread(target=producer)
        th.start()
        th.join()
        offered = n_req / max(time.perf_counter() - t0, 1e-9)
        for tk in tickets:
            st, _ = srv.wait(tk)
            assert st == 0
        for tk in fit_tickets:
            srv.wait(tk)
        s = srv.stats()
        srv.close()
        return {"p50_ms": round(s["p50_ms"], 3), "p99_ms": round(s["p99_ms"], 3),
                "rows_per_s": round(s["rows_per_s"], 1), "requests": s["predict_requests"],
                "requests_per_s_submitted": round(offered, 1), "launches": s["predict_launches"],
                "fit_jobs": s["fit_jobs"], "hbm_peak_gb": round(s["hbm_peak"] / 1e9, 2),
                "fit_mem_measured_over_admitted": round(s["fit_mem_ratio_max"], 3)}

    result = {"forest_trees": forest.n_trees, "forest_nodes": forest.n_nodes, "forest_fit_s": round(t_fit, 2),
              "stream": "Poisson arrivals for %.0fs per rate, predict rows log-uniform 1..65536, one fit job "
                        "(1k-200k rows, 10-50 trees) every 2 s; coalesced = short (<= 4,096 rows) and long "
                        "request lanes, per_request = one launch per request on one lane" % dur,
              "sweep": []}
    for qps in args.serving_qps:
        e = {"qps": qps}
        for coalesce in (True, False):
            key = "coalesced" if coalesce else "per_request"
            e[key] = stream(qps, 65536, True, coalesce, 11)
            log("serving %5.0f req/s %-11s: p50 %.2f ms p99 %.2f ms, %.0f rows/s, %d launches for %d requests" %
                (qps, key, e[key]["p50_ms"], e[key]["p99_ms"], e[key]["rows_per_s"], e[key]["launches"],
                 e[key]["requests"]))
        e["rows_per_s_ratio"] = round(e["coalesced"]["rows_per_s"] / max(e["per_request"]["rows_per_s"], 1e-9), 2)
        result["sweep"].append(e)
    def burst(n_req, largest, coalesce, seed, per_call=1000):
        # every request submitted at once (a client handing over requests it
        # has gathered, per_call of them per submit call): the completion
        # rate is the serving capacity of the mode
        rng = np.random.default_rng(seed)
        sizes = np.minimum(np.exp(rng.uniform(0, np.log(largest), size=n_req)).astype(int), largest)
        starts = rng.integers(0, len(Xt) - largest, size=n_req)
        srv = lib.Server(coalesce=coalesce, fit_workers=1)
        fid = srv.add_forest(forest)
        # the gathered requests lie back to back in one array
        Xall = np.concatenate([Xt[starts[i]:starts[i] + sizes[i]] for i in range(n_req)])
        first = np.concatenate(([0], np.cumsum(sizes)))
        tickets = []
        t0 = time.perf_counter()
        for c in range(0, n_req, per_call):
            e = min(c + per_call, n_req)
            t, _ = srv.submit_rows(fid, Xall[first[c]:first[e]], sizes[c:e])
            tickets.append(t)
        t_sub = time.perf_counter() - t0
        st = srv.wait_many(np.concatenate(tickets))
        assert (st == 0).all()
        t_all = time.perf_counter() - t0
        s = srv.stats()
        srv.close()
        return {"requests": n_req, "rows": int(sizes.sum()), "submit_s": round(t_sub, 3), "wall_s": round(t_all, 3),
                "requests_per_s": round(n_req / t_all, 1), "rows_per_s": round(s["rows_per_s"], 1),
                "p50_ms": round(s["p50_ms"], 3), "p99_ms": round(s["p99_ms"], 3), "launches": s["predict_launches"]}

    if args.serving_burst > 0:
        e = {"request_rows": "log-uniform 1..64",
             "submit": "all at once, 1,000 requests per submit call (rows back to back), one wait for all"}
        for coalesce in (True, False):
            key = "coalesced" if coalesce else "per_request"
            e[key] = burst(args.serving_burst, 64, coalesce, 13)
            log("serving burst %d small requests %-11s: %.0f req/s, %.0f rows/s, %d launches" %
                (args.serving_burst, key, e[key]["requests_per_s"], e[key]["rows_per_s"], e[key]["launches"]))
        e["rows_per_s_ratio"] = round(e["coalesced"]["rows_per_s"] / max(e["per_request"]["rows_per_s"], 1e-9), 2)
        result["saturation_small_requests"] = e
    if args.serving_small_qps > 0:
        e = {"qps": args.serving_small_qps, "request_rows": "log-uniform 1..64, no fit jobs"}
        for coalesce in (True, False):
            key = "coalesced" if coalesce else "per_request"
            e[key] = stream(args.serving_small_qps, 64, False, coalesce, 12)
            log("serving small %5.0f req/s %-11s: p50 %.2f ms p99 %.2f ms, %.0f rows/s, %d launches" %
                (args.serving_small_qps, key, e[key]["p50_ms"], e[key]["p99_ms"], e[key]["rows_per_s"],
                 e[key]["launches"]))
        e["rows_per_s_ratio"] = round(e["coalesced"]["rows_per_s"] / max(e["per_request"]["rows_per_s"], 1e-9), 2)
        result["small_requests"] = e
    return result


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "pmc_traffic_fit_covtype_581k_x54_50trees.md")


def short_kernel(name):
    """rocprofv3 kernel name -> the name bench.py times it under."""
    base = name.split("(")[0].replace("void ", "").strip()
    if base.startswith("cta::"):
        base = base[5:]
    ident = base.split("<")[0]
    return ident if base.startswith("k_") else base


def pmc_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE, calibrated by a plain 1 GiB copy)."""
    import re

    if not os.path.exists(TRAFFIC_SUMMARY):
        return None, None
    for line in open(TRAFFIC_SUMMARY):
        m = re.match(r"- (\w+): traffic_bytes_per_launch=(\d+)", line.strip())
        if m and m.group(1) == kernel:
            return int(m.group(2)), os.path.relpath(TRAFFIC_SUMMARY, ROOT)
    return None, None


def pmc_summary(fetch_csv, write_csv, l2_csv, out):
    """Per-kernel FETCH_SIZE / WRITE_SIZE (KiB per dispatch) of two rocprofv3
    --pmc runs of `bench.py --calibrate-copy`, scaled by what the counters
    read for the 1 GiB device-to-device copy in the same runs."""
    import csv
    from collections import defaultdict

    def load(path, counter):
        tot, cnt, copies = defaultdict(float), defaultdict(int), []
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != counter:
                continue
            v = float(r["Counter_Value"])
            k = short_kernel(r["Kernel_Name"])
            tot[k] += v
            cnt[k] += 1
            if "copyBuffer" in r["Kernel_Name"]:
                copies.append(v)
        return tot, cnt, copies

    ft, fc, fcopy = load(fetch_csv, "FETCH_SIZE")
    wt, wc, wcopy = load(write_csv, "WRITE_SIZE")
    gib_kib = float(1 << 20)
    wcal = max(wcopy)  # the 1 GiB copy is the largest copy dispatch
    fcal = max(fcopy)
    ffac, wfac = gib_kib / fcal, gib_kib / wcal
    l2 = {}
    if l2_csv:
        hit, miss = defaultdict(float), defaultdict(float)
        for r in csv.DictReader(open(l2_csv)):
            k = short_kernel(r["Kernel_Name"])
            if r["Counter_Name"].startswith("TCC_HIT"):
                hit[k] += float(r["Counter_Value"])
            elif r["Counter_Name"].startswith("TCC_MISS"):
                miss[k] += float(r["Counter_Value"])
        l2 = {k: hit[k] / (hit[k] + miss[k]) for k in hit if hit[k] + miss[k] > 0}
    lines = ["| kernel | dispatches | FETCH_SIZE KiB/disp | WRITE_SIZE KiB/disp | HBM bytes/launch (calibrated) | L2 hit |",
             "|---|---:|---:|---:|---:|---:|"]
    rows = []
    for k in sorted(ft, key=lambda k: -(ft[k] * ffac + wt.get(k, 0) * wfac)):
        if k.startswith("__amd") or fc[k] == 0:
            continue
        f = ft[k] / fc[k]
        w = wt.get(k, 0.0) / max(1, wc.get(k, 1))
        t = (f * ffac + w * wfac) * 1024
        rows.append((k, fc[k], f, w, t))
        lines.append("| %s | %d | %.1f | %.1f | %.4g | %s |" % (k, fc[k], f, w, t, "%.3f" % l2[k] if k in l2 else ""))
    out.write("Calibration: the 1 GiB copy read %.0f KiB by FETCH_SIZE and wrote %.0f KiB by WRITE_SIZE, so\n"
              "fetch factor %.3f and write factor %.3f convert the counters to bytes.\n\n" % (fcal, wcal, ffac, wfac))
    out.write("\n".join(lines) + "\n\nMachine-readable (bench.py reads these):\n\n")
    for k, n, f, w, t in rows:
        out.write("- %s: traffic_bytes_per_launch=%d fetch_kib=%.1f write_kib=%.1f dispatches=%d\n" % (k, t, f, w, n))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-trees", type=int, default=16, help="trees in the bounded CPU-oracle sample")
    ap.add_argument("--cpu-threads", type=int, default=0, help="host threads growing them (0: every core)")
    ap.add_argument("--cpu-trees-1core", type=int, default=4, help="trees of the one-core oracle figure (0: skip)")
    ap.add_argument("--per-node-trees", type=int, default=1, help="trees in the per-node GPU sample (0: skip)")
    ap.add_argument("--sklearn-trees", type=int, default=8, help="trees in the scikit-learn sample (0: skip)")
    ap.add_argument("--no-profile", action="store_true", help="skip per-kernel HIP-event timing")
    ap.add_argument("--serving-seconds", type=float, default=3.0, help="length of each serving stream (0: skip)")
    ap.add_argument("--serving-trees", type=int, default=500, help="trees of the resident serving forest")
    ap.add_argument("--serving-qps", type=float, nargs="+", default=[500.0, 1000.0, 2000.0, 4000.0],
                    help="predict request rates of the serving sweep")
    ap.add_argument("--serving-small-qps", type=float, default=20000.0,
                    help="request rate of the small-request stream (0: skip)")
    ap.add_argument("--serving-burst", type=int, default=20000,
                    help="small requests submitted at once to measure each mode's capacity (0: skip)")
    ap.add_argument("--workload-steps", type=int, default=3, help="timed fits per other workload (0: skip them)")
    ap.add_argument("--tune-probe-trees", type=int, default=8,
                    help="trees per autotuner probe on the 200,000 x 256 workload (0: no sweep)")
    ap.add_argument("--facade-steps", type=int, default=3, help="timed facade fits from host arrays (0: skip)")
    ap.add_argument("--pmc-summary", nargs="+", metavar="CSV",
                    help="FETCH_SIZE csv, WRITE_SIZE csv [, TCC_HIT/TCC_MISS csv]: print the traffic summary and exit")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one plain 1 GiB device-to-device copy (calibrates FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.pmc_summary:
        c = args.pmc_summary
        pmc_summary(c[0], c[1], c[2] if len(c) > 2 else None, sys.stdout)
        return
    if args.gpus != 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path (trees would shard across GPUs "
              "with one gather of the node tables); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    from cudatree_amd import _lib as lib
    from cudatree_amd import datasets

    t0 = time.time()
    X, y, Xt, yt = datasets.covtype_like()
    N, F = X.shape
    K = int(y.max()) + 1
    log("data %.1fs: N=%d F=%d K=%d" % (time.time() - t0, N, F, K))

    ctx = lib.default_context(0)
    dev = torch.device("cuda:0")
    Xd = torch.from_numpy(X).to(dev)
    yd = torch.from_numpy(y.astype(np.uint32).view(np.int32)).to(dev)
    Xtd = torch.from_numpy(Xt).to(dev)
    torch.cuda.synchronize()

    for w in range(args.warmup):
        t = time.time()
        forest, arrays = fit_once(ctx, lib, Xd, yd, N, F, K)
        log("warmup %d: %.3fs, %d nodes" % (w, time.time() - t, len(arrays["feature"])))
        del forest

    def timed_steps():
        ctx.synchronize()
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        for s in range(args.steps):
            forest, _ = fit_once(ctx, lib, Xd, yd, N, F, K, export=False)
        ctx.synchronize()
        torch.cuda.synchronize()
        t = (time.perf_counter() - t_start) / args.steps
        return t, forest, forest.export()

    # the fit as a caller runs it: no per-kernel events
    t_fit, forest, arrays = timed_steps()
    log("fit: %.4f s per 50 trees (%d nodes)" % (t_fit, len(arrays["feature"])))
    # the same steps again with HIP events around every kernel group, for
    # the per-kernel times and the roofline
    t_prof = None
    ctx.reset_stats()
    if not args.no_profile:
        ctx.set_profiling(True)
        del forest
        t_prof, forest, arrays = timed_steps()
        ctx.set_profiling(False)
        log("fit with per-kernel events: %.4f s per 50 trees" % t_prof)
    stats = ctx.kernel_stats()
    for k, v in sorted(stats.items(), key=lambda kv: -kv[1]["ms"]):
        gbs = v["alg_bytes"] / (v["ms"] * 1e-3) / 1e9 if v["ms"] > 0 and v["alg_bytes"] else 0.0
        log("  %-16s %8d launches %10.3f ms  %8.1f GB/s alg" % (k, v["launches"], v["ms"] / args.steps, gbs))

    # roofline of the dominant HBM kernel, HIP events over the timed steps
    roof = None
    hbm_kernels = [k for k in ("k_resh", "k_scan", "k_subtree", "k_rl_count")
                   if k in stats and stats[k]["ms"] > 0]
    if hbm_kernels:
        dom = max(hbm_kernels, key=lambda k: stats[k]["ms"])
        per_kernel = {}
        for k in hbm_kernels:
            s = stats[k]
            per_launch_bytes = s["alg_bytes"] / s["launches"]
            per_launch_s = s["ms"] * 1e-3 / s["launches"]
            per_kernel[k] = {"achieved": round(per_launch_bytes / per_launch_s / 1e9, 1),
                             "ms_per_step": round(s["ms"] / args.steps, 3), "launches": s["launches"],
                             "avg_launch_us": round(per_launch_s * 1e6, 2),
                             "alg_bytes_per_launch": int(per_launch_bytes)}
        d = per_kernel[dom]
        traffic, tsrc = pmc_traffic(dom)
        # split scan + reshuffle kernels together (their algorithmic bytes over
        # their summed time), the figure the >= 50 % of peak target is set on
        group = [k for k in ("k_scan", "k_scan_hist", "k_rl_count", "k_resh", "k_fill",
                             "k_subtree") if k in stats]
        g_bytes = sum(stats[k]["alg_bytes"] for k in group)
        g_s = sum(stats[k]["ms"] for k in group) * 1e-3
        combined = {"kernels": group, "achieved": round(g_bytes / g_s / 1e9, 1) if g_s else None,
                    "frac": round(g_bytes / g_s / 1e9 / HBM_PEAK_GBS, 4) if g_s else None}
        roof = {"bound": "hbm", "kernel": dom, "achieved": d["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(d["achieved"] / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": tsrc,
                "launches": d["launches"], "avg_launch_us": d["avg_launch_us"],
                "alg_bytes_per_launch": d["alg_bytes_per_launch"], "scan_and_reshuffle": combined,
                "kernels": per_kernel}

    # ---- predict: 100,000 held-out rows resident in HBM
    R = Xt.shape[0]
    labels = torch.empty(R, dtype=torch.int32, device=dev)
    import ctypes

    def pred(nrows):
        lib.check(lib.load().cta_predict_device(ctx.h, forest.h, ctypes.c_void_p(Xtd.data_ptr()), nrows, F,
                                                ctypes.c_void_p(labels.data_ptr()), None))

    pred(R)
    t = time.perf_counter()
    reps = 10
    for _ in range(reps):
        pred(R)
    t_pred = (time.perf_counter() - t) / reps
    got = labels.cpu().numpy().astype(np.uint32)
    acc = float(np.mean(got == yt))
    # latency per request, log-uniform 1..65,536 rows, one request per launch
    rng = np.random.default_rng(7)
    sizes = np.exp(rng.uniform(0, np.log(65536), size=200)).astype(int)
    lat = []
    for n in sizes:
        t = time.perf_counter()
        pred(int(n))
        lat.append((time.perf_counter() - t) * 1e3)
    lat = np.array(lat)
    predict = {"rows": R, "rows_per_s": round(R / t_pred, 1), "ms_per_batch": round(t_pred * 1e3, 3),
               "accuracy": round(acc, 4), "request_rows": "log-uniform 1..65536, 200 requests",
               "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3)}
    log("predict: %.0f rows/s, acc %.4f, p50 %.3f ms p99 %.3f ms" % (predict["rows_per_s"], acc, predict["p50_ms"],
                                                                    predict["p99_ms"]))

    # ---- same kernels, one node per launch sequence (bounded sample)
    per_node = None
    if args.per_node_trees > 0:
        nt = args.per_node_trees
        t = time.perf_counter()
        f2, a2 = fit_once(ctx, lib, Xd, yd, N, F, K, mode="per_node", n_trees=nt)
        dt = time.perf_counter() - t
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import helpers  # tests/helpers.py: exact forest comparison

        same = helpers.forest_diff(helpers.forest_slice(arrays, 0, nt), a2) is None
        per_node = {"sample": "%d of %d trees, per-node launches, scaled x%g" % (nt, N_TREES, N_TREES / nt),
                    "s_per_50_trees": round(dt * N_TREES / nt, 3), "speedup_batched": round(dt * N_TREES / nt / t_fit, 2),
                    "identical_trees": bool(same)}
        log("per-node: %.2fs for %d tree(s) -> %.1fs/50, batched speedup %.1fx" % (dt, nt, dt * N_TREES / nt,
                                                                                 dt * N_TREES / nt / t_fit))
        del f2

    # ---- CPU baselines (bounded samples on the host cores)
    cpu = None
    if args.cpu_trees > 0:
        import oracle

        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import helpers  # tests/helpers.py: exact forest comparison

        o = oracle.get()
        cores = len(os.sched_getaffinity(0))
        threads = args.cpu_threads if args.cpu_threads > 0 else cores
        nt = args.cpu_trees
        ids = o.bootstrap_forest_ids(SEED, nt, N)
        # the oracle's trees in parallel on the host cores (one tree per thread
        # at a time), then a few trees on one core
        t = time.perf_counter()
        fo = o.fit(X, y, K, nt, MAX_FEATURES, boot_ids=ids, seed=SEED, threads=threads)
        dt = time.perf_counter() - t
        # every field of the first nt trees: thresholds bit for bit, sample
        # counts, depths and the tree boundaries
        same = helpers.forest_diff(helpers.forest_slice(arrays, 0, nt), fo) is None
        cpu = {"value": round(dt * N_TREES / nt, 2), "unit": "s/50 trees", "cores": threads, "kind": "port",
               "sample": "%d of %d trees of the same workload grown by %d host threads, scaled x%.3g" %
                         (nt, N_TREES, threads, N_TREES / nt),
               "identical_trees": bool(same)}
        log("cpu oracle: %.2fs for %d trees on %d threads -> %.1fs/50 (identical: %s)" %
            (dt, nt, threads, dt * N_TREES / nt, same))
        if args.cpu_trees_1core > 0:
            n1 = args.cpu_trees_1core
            t = time.perf_counter()
            o.fit(X, y, K, n1, MAX_FEATURES, boot_ids=ids[:n1], seed=SEED)
            d1 = time.perf_counter() - t
            cpu["one_core"] = {"value": round(d1 * N_TREES / n1, 2), "unit": "s/50 trees",
                               "sample": "%d trees on one core, scaled x%.3g" % (n1, N_TREES / n1)}
            log("cpu oracle one core: %.2fs for %d trees -> %.1fs/50" % (d1, n1, d1 * N_TREES / n1))
    sk = None
    if args.sklearn_trees > 0:
        from sklearn.ensemble import RandomForestClassifier

        cores = len(os.sched_getaffinity(0))
        nt = args.sklearn_trees
        t = time.perf_counter()
        rf = RandomForestClassifier(n_estimators=nt, max_features=MAX_FEATURES, n_jobs=-1, random_state=0).fit(X, y)
        dt = time.perf_counter() - t
        t = time.perf_counter()
        rf.predict(Xt)
        dp = time.perf_counter() - t
        sk = {"fit_s_per_50_trees": round(dt * N_TREES / nt, 2), "cores": cores,
              "sample": "%d trees scaled x%.3g" % (nt, N_TREES / nt),
              "predict_rows_per_s_%d_trees" % nt: round(R / dp, 1)}
        log("sklearn: %.2fs for %d trees on %d cores -> %.1fs/50" % (dt, nt, cores, dt * N_TREES / nt))

    if args.calibrate_copy:
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: 1 GiB read + 1 GiB written")
        del src, dst

    # ---- the facade as an API caller runs it: host X and labels to a forest
    facade = None
    if args.facade_steps > 0:
        import cudatree_amd

        times = []
        for k in range(args.facade_steps + 1):
            t = time.perf_counter()
            rf = cudatree_amd.RandomForestClassifier(n_estimators=N_TREES, max_features=MAX_FEATURES,
                                                     random_state=SEED).fit(X, y)
            dt = time.perf_counter() - t
            if k:
                times.append(dt)
            del rf
        facade = {"s_per_fit": round(float(np.median(times)), 4), "steps": args.facade_steps,
                  "what": "RandomForestClassifier(n_estimators=50, max_features=7).fit(X, y) from host numpy "
                          "arrays: label compaction, upload, presort, fit, forest resident for predict"}
        log("facade fit from host arrays: %.4f s" % facade["s_per_fit"])

    workloads = None
    if args.workload_steps > 0:
        del forest
        ctx.release_workspace()
        workloads = workload_bench(lib, ctx, args, torch, dev)

    serving = None
    if args.serving_seconds > 0:
        serving = serving_bench(lib, X, y, Xt, args)

    out = {"metric": METRIC, "value": round(t_fit, 4), "unit": "s/50 trees", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(t_fit * 1e3, 2),
           "ms_per_step_with_kernel_events": round(t_prof * 1e3, 2) if t_prof else None, "higher_is_better": False,
           "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic",
           "config": {"workload": "covtype-shaped 581,012 x 54 float32, 7 classes; 50 trees, bootstrap, "
                                  "max_features=7, full depth; predict 100,000 held-out rows",
                      "n_samples": N, "n_features": F, "n_classes": K, "n_trees": N_TREES,
                      "nodes": int(len(arrays["feature"])), "mode": "batched"},
           "roofline": roof, "cpu_baseline": cpu, "facade_fit_s": facade["s_per_fit"] if facade else None,
           "facade": facade, "predict": predict, "per_node_gpu": per_node, "sklearn": sk, "workloads": workloads,
           "serving": serving}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
