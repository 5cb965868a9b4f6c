"""Benchmark: epochs/s of 2-D FIt-SNE at N = 1,000,000 on one MI355X.

A step is one epoch of the FIt-SNE loop (recentring, repulsion by FFT interpolation beside the
attraction over the column-tiled P, then the gains/momentum update) over the whole embedding.  The workload is the
1M-point configuration: 50-component Gaussian mixture in 50-D (seed 2), perplexity 30 (k = 90),
grid >= 50 boxes per axis.  Affinities (kNN + bandwidth search + symmetrisation) are built on the
GPU before the timed region and reported separately; data, P and the optimiser state are resident
in HBM when timing starts.

Usage: python bench.py [--gpus 1] [--steps 750] [--warmup 10] [--workload c3|c2|c1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "epochs/s at N=1M 2D FIt-SNE; batched jobs/s + p50/p99 latency vs plain"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "c3": dict(n=1_000_000, name="FIt-SNE 2D, N=1,000,000, D_in=50, 50-component GMM "
               "(centres N(0,36I), sigma 1, Dirichlet(1) weights, seed 2), perplexity 30 (k=90), "
               "grid >= 50 boxes/axis", min_boxes=50, method="fitsne", epochs=750),
    "c2": dict(n=70_000, name="FIt-SNE 2D, N=70,000, D_in=20, 10-component GMM (centres "
               "N(0,16I), axis std 1/sqrt(i+1), seed 1), perplexity 30", min_boxes=10,
               method="fitsne", epochs=750),
    "c1": dict(n=5_000, name="Barnes-Hut 2D theta 0.5, N=5,000, D_in=50, 10-component GMM "
               "(centres N(0,25I), sigma 1, equal weights, seed 0), perplexity 30 (k=90)",
               min_boxes=10, method="bh", epochs=1000),
    "c4": dict(name="batched: independent jobs seeded by job id, N log-uniform [500, 50,000], "
               "D_in in {16,32,64}, 5-20 component GMM (centres N(0,25I), sigma 1), D_out 2, "
               "FIt-SNE (even ids) / Barnes-Hut theta 0.5 (odd ids), perplexity 30"),
    "c5": dict(name="serving: Poisson arrivals, requests seeded by id, N log-uniform "
               "[1,000, 200,000], D_in in {20,50}, D_out in {1,2}, 10-component GMM, FIt-SNE, "
               "perplexity 30"),
}


def bench_batched(args):
    """Batched jobs/s: every job in one batch (affinity builds, then all epochs advanced by
    shared launches), against plain one-job-at-a-time fits of every `stride`-th job."""
    import bhtsne_amd as bt
    from bhtsne_amd import datagen
    njobs, epochs, stride = args.jobs, args.job_epochs, args.plain_stride
    specs = [datagen.config_c4_job(j) for j in range(njobs)]
    ns = [x.shape[0] for x, _ in specs]
    log(f"{njobs} jobs, N {min(ns)}..{max(ns)} (sum {sum(ns)}), {epochs} epochs each")
    ctx = bt.Context(0)

    def make(j):
        x, _ = specs[j]
        t = bt.TSNE(x, ctx=ctx).perplexity(30.0).epochs(epochs)
        return t.initial_embedding(datagen.seed_embedding(x.shape[0], 2, j))

    # warm the kernels (first launches load the code objects)
    graph = not getattr(args, "no_graph", False)
    w = bt.Batch(ctx, graph=graph)
    w.add(make(0), specs[0][1], 0.5)
    w.add(make(1), specs[1][1], 0.5)
    w.run()
    w.close()

    # admission: the jobs' affinities and optimiser state are prepared by `preparers` host threads
    # (each job on its own stream, as the request loop does), then joined in job order
    from concurrent.futures import ThreadPoolExecutor

    def prepared(j):
        t = make(j)
        t.prepare(specs[j][1], 0.5)
        return t

    t0 = time.perf_counter()
    b = bt.Batch(ctx, graph=graph)
    with ThreadPoolExecutor(max_workers=args.preparers) as ex:
        jobs = list(ex.map(prepared, range(njobs)))
    for j in range(njobs):
        b.add(jobs[j], specs[j][1], 0.5)
    t_admit = time.perf_counter() - t0
    done = b.run()
    t_batch = time.perf_counter() - t0
    assert len(done) == njobs
    embs = {j: jobs[j].embedding() for j in range(0, njobs, stride)}
    kls = {j: jobs[j].kl_divergence() for j in range(0, njobs, stride)}
    b.close()
    log(f"batched: {njobs} jobs in {t_batch:.2f}s (admission incl. affinities {t_admit:.2f}s, "
        f"epochs {t_batch - t_admit:.2f}s) -> {njobs / t_batch:.2f} jobs/s")
    del jobs, done

    # plain: each sampled job fitted alone
    plain_t, same, max_kl = 0.0, 0, 0.0
    sample = list(range(0, njobs, stride))
    for j in sample:
        t = make(j)
        s0 = time.perf_counter()
        if specs[j][1] == "fitsne":
            t.fit_sne()
        else:
            t.barnes_hut(0.5)
        t.synchronize()
        plain_t += time.perf_counter() - s0
        same += int(np.array_equal(t.embedding(), embs[j]))
        max_kl = max(max_kl, abs(t.kl_divergence() - kls[j]) / abs(kls[j]))
    plain_rate = len(sample) / plain_t
    log(f"plain: {len(sample)} sampled jobs in {plain_t:.2f}s -> {plain_rate:.2f} jobs/s; "
        f"{same}/{len(sample)} embeddings bit-identical to the batch, max KL rel diff {max_kl:.1e}")
    rate = njobs / t_batch
    return {"metric": f"batched jobs/s ({njobs} mixed FIt-SNE / Barnes-Hut jobs, "
                      f"{epochs} epochs each)",
            "value": round(rate, 3), "unit": "jobs/s", "n_gpus": 1, "steps": njobs,
            "warmup": 2, "ms_per_step": round(1e3 * t_batch / njobs, 3),
            "higher_is_better": True, "scaling": "strong", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic",
            "config": {"workload": WORKLOADS["c4"]["name"], "jobs": njobs, "epochs": epochs,
                       "sum_n": sum(ns), "parallelism": "single GPU"},
            "batched": {"jobs_per_s": round(rate, 3), "seconds": round(t_batch, 3),
                        "admission_s": round(t_admit, 3), "preparers": args.preparers},
            "plain": {"jobs_per_s": round(plain_rate, 3), "sampled_jobs": len(sample),
                      "seconds": round(plain_t, 3)},
            "vs_plain": round(rate / plain_rate, 3),
            "parity": {"bit_identical": f"{same}/{len(sample)}", "max_kl_rel_diff": max_kl}}


def bench_serving(args):
    """Serving: a Poisson stream of requests through the continuous-batching loop, then the same
    stream served one plain fit at a time; jobs/s, p50 / p99 latency, per-job equality."""
    import bhtsne_amd as bt
    from bhtsne_amd import datagen, serving
    count, rate, epochs = args.requests, args.rate, args.job_epochs
    arrivals = datagen.poisson_arrivals(count, rate, 0)
    reqs = []
    for r in range(count):
        x, d_out = datagen.config_c5_request(r)
        reqs.append(serving.Request(r, x, d_out=d_out, epochs=epochs, arrival=float(arrivals[r]),
                                    y0=datagen.seed_embedding(x.shape[0], d_out, r)))
    ns = [q.x.shape[0] for q in reqs]
    log(f"{count} requests at {rate}/s, N {min(ns)}..{max(ns)}, {epochs} epochs each")
    ctx = bt.Context(0)
    # warm-up request
    serving.serve([serving.Request(-1, reqs[0].x[:2000], epochs=20)], ctx, want_kl=False)
    res_b = serving.serve(reqs, ctx, step_epochs=args.step_epochs,
                          graph=not getattr(args, "no_graph", False))
    sb = serving.latency_summary(res_b)
    log(f"batched: {sb}; loop {serving.serve.last_stats}")
    res_p = serving.serve_plain(reqs, ctx)
    sp = serving.latency_summary(res_p)
    log(f"plain: {sp}")
    same = sum(int(np.array_equal(a.embedding, c.embedding)) for a, c in zip(res_b, res_p))
    kl_diff = max(abs(a.kl - c.kl) / abs(c.kl) for a, c in zip(res_b, res_p))
    return {"metric": "served jobs/s with p50/p99 latency (Poisson arrivals)",
            "value": round(sb["jobs_per_s"], 3), "unit": "jobs/s", "n_gpus": 1,
            "steps": count, "warmup": 1, "ms_per_step": round(1e3 * sb["span_s"] / count, 3),
            "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic",
            "config": {"workload": WORKLOADS["c5"]["name"], "requests": count,
                       "rate_per_s": rate, "epochs": epochs, "parallelism": "single GPU"},
            "batched": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in sb.items()},
            "plain": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in sp.items()},
            "vs_plain": round(sb["jobs_per_s"] / sp["jobs_per_s"], 3),
            "parity": {"bit_identical": f"{same}/{count}", "max_kl_rel_diff": kl_diff}}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_data(workload):
    from bhtsne_amd import datagen
    if workload == "c3":
        return datagen.config_c3()[0]
    if workload == "c2":
        return datagen.config_c2()[0]
    return datagen.config_c1()[0]


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` at the 1M-point workload, as recorded from the FETCH_SIZE /
    WRITE_SIZE counter passes (scaled by the copy calibration) in the profile summary."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "k_attract_c3_hbm_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(f"traffic_bytes_per_launch {kernel}:"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def cpu_baseline(aff, y_seed, state_late, late_epoch, min_boxes, n_epochs=3, method="fitsne"):
    """The oracle (a C restatement of the reference's f32 loop, OpenMP over rows as rayon is)
    timed on this host: n_epochs from the seed state (epoch 0) and n_epochs from the GPU run's
    state at epoch `late_epoch` (after stop-lying), averaged per epoch."""
    import oracle
    threads = oracle.set_threads(0)
    rows = np.ascontiguousarray(aff.rows, np.uint64)
    cols = np.ascontiguousarray(aff.columns, np.uint32)
    n = y_seed.shape[0]
    lr = oracle.auto_learning_rate(n, 12.0)
    times = []
    for label, (y, uy, g), vals_scale, e0 in (
            ("early", (y_seed.copy(), np.zeros_like(y_seed), np.ones_like(y_seed)), 12.0, 0),
            ("late", tuple(a.copy() for a in state_late), 1.0, late_epoch)):
        vals = (np.ascontiguousarray(aff.values, np.float32) * np.float32(vals_scale))
        t0 = time.perf_counter()
        oracle.run_epochs(y, uy, g, rows, cols, vals, method=method, min_boxes=min_boxes,
                          theta=0.5, learning_rate=lr, e0=e0, e1=e0 + n_epochs,
                          momentum=0.5 if label == "early" else 0.8)
        dt = time.perf_counter() - t0
        times.append(dt / n_epochs)
        log(f"cpu baseline {label} (epochs {e0}..{e0 + n_epochs}): {dt / n_epochs:.3f} s/epoch")
    per_epoch = float(np.mean(times))
    fft = ("; its FFT is a recursive mixed-radix complex transform in double precision with an "
           "allocation per line, slower than the reference's rustfft/realfft, so this rate is a "
           "lower bound of the reference's and the GPU/CPU ratio an upper bound"
           if method == "fitsne" else "")
    return {"value": 1.0 / per_epoch, "unit": "epochs/s", "cores": int(threads), "kind": "port",
            "sample": f"oracle (C restatement of the reference loop, "
                      f"{'grid >= %d boxes' % min_boxes if method == 'fitsne' else 'theta 0.5'})"
                      f" on this host: {n_epochs} epochs from the seed state (epoch 0) + "
                      f"{n_epochs} epochs from the GPU run's state at epoch {late_epoch}, mean "
                      f"s/epoch (early {times[0]:.3f}, late {times[1]:.3f}){fft}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=750)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-epochs", type=int, default=40,
                    help="CPU baseline: epochs timed from the seed state and from the final state")
    ap.add_argument("--no-batched", action="store_true",
                    help="leave out the batched and serving summaries of the default line")
    ap.add_argument("--jobs", type=int, default=256, help="batched workload: number of jobs")
    ap.add_argument("--preparers", type=int, default=4,
                    help="batched workload: host threads preparing the jobs before they join")
    ap.add_argument("--job-epochs", type=int, default=1000,
                    help="batched / serving workloads: epochs per job")
    ap.add_argument("--plain-stride", type=int, default=4,
                    help="batched workload: fit every k-th job alone for the plain rate")
    ap.add_argument("--requests", type=int, default=48, help="serving workload: requests")
    ap.add_argument("--rate", type=float, default=400.0,
                    help="serving workload: arrivals per s (saturating: the loop's capacity)")
    ap.add_argument("--serial-epoch", action="store_true",
                    help="run each epoch's kernels in order on one stream (per-kernel profiles)")
    ap.add_argument("--no-graph", action="store_true",
                    help="batched and serving workloads: plain launches instead of the captured "
                         "epoch graph")
    ap.add_argument("--step-epochs", type=int, default=50,
                    help="serving workload: epochs between admissions")
    ap.add_argument("--short", action="store_true",
                    help="only the timed epochs (for profiler runs): no error report, per-phase "
                         "pass, full run, reference-grid run, CPU baseline or batched lines")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB five times (the "
                         "known-size kernel that scales the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_probe > 0:
        import torch
        src = torch.ones(args.copy_probe << 18, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": src.numel() * 4, "launches": 5}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs every job on a single MI355X "
              f"(--gpus {args.gpus} is not supported)", file=sys.stderr)
        sys.exit(2)

    if args.workload in ("c4", "c5"):
        out = bench_batched(args) if args.workload == "c4" else bench_serving(args)
        print(json.dumps(out), flush=True)
        return

    import bhtsne_amd as bt
    from bhtsne_amd import datagen, metrics

    wl = WORKLOADS[args.workload]
    method = wl["method"]
    total_epochs = wl["epochs"]
    steps = args.steps
    seed = {"c1": 0, "c2": 1, "c3": 2}[args.workload]
    t0 = time.perf_counter()
    x = make_data(args.workload)
    n = x.shape[0]
    log(f"data {x.shape} generated in {time.perf_counter() - t0:.1f}s "
        f"sha256={datagen.sha256(x)[:16]}")
    y_seed = datagen.seed_embedding(n, 2, seed)

    t = bt.TSNE(x).perplexity(30.0).epochs(total_epochs).min_boxes(wl["min_boxes"])
    if args.serial_epoch:
        t._set(flags=2)  # BHT_FLAG_SERIAL_EPOCH: repulsion and attraction in order, one stream
    # affinities: kNN + bandwidth search + symmetrisation on the GPU (not in the timed region)
    t0 = time.perf_counter()
    t.initial_embedding(y_seed).prepare(method, 0.5)
    t.synchronize()
    t_aff = time.perf_counter() - t0
    aff = t.affinities()
    nnz = int(aff.columns.size)
    log(f"affinities built in {t_aff:.2f}s, nnz={nnz}")

    # warmup run (discarded), then a fresh run from the seed state, timed with no events
    if args.warmup > 0:
        t.step(args.warmup)
        t.synchronize()
    t.initial_embedding(y_seed).prepare(method, 0.5)
    t.synchronize()
    start = time.perf_counter()
    t.step(steps)
    t.synchronize()
    elapsed = time.perf_counter() - start
    eps = steps / elapsed
    form = t.epoch_form()
    grid = t.grid_info() if method == "fitsne" else None
    if args.short:
        log(f"{steps} epochs in {elapsed:.4f}s -> {eps:.1f} epochs/s; grid {grid}; "
            f"epoch form {form}")
        print(json.dumps({"metric": METRIC, "value": round(eps, 3), "unit": "epochs/s",
                          "steps": steps, "short": True}), flush=True)
        return
    # the state after the timed epochs against the exact FP64 gradient (all-pairs repulsion and Z
    # on the device, edge terms in FP64 on a 5,000-row sample)
    t0 = time.perf_counter()
    errs = metrics.error_report(t, sample=5000, seed=0)
    log(f"vs FP64 exact (5k rows, epoch {steps}): gradient rel L2 {errs['gradient']:.3e}, "
        f"repulsion F-/Z rel L2 {errs['repulsion']:.3e}, Z rel {errs['z']:.2e} "
        f"({time.perf_counter() - t0:.1f}s)")
    t.finish()
    kl_steps = t.kl_divergence_pair()
    log(f"{steps} epochs in {elapsed:.4f}s -> {eps:.1f} epochs/s; grid {grid}; KL {kl_steps}")

    # per-phase device time: the same epochs again from the seed state with HIP events on the
    # job's streams, then the run continued (untimed) to its full length for the final KL and
    # the late state of the CPU baseline
    t.initial_embedding(y_seed).prepare(method, 0.5)
    t.set_timing(True)
    t.step(steps)
    rep_ms, att_ms, timed = t.timing()
    t.set_timing(False)
    late_epoch = min(600, total_epochs)
    if late_epoch > steps:
        t.step(late_epoch - steps)
    state = t.get_state()
    t.step(total_epochs - max(late_epoch, steps))
    t.synchronize()
    final_grid = t.grid_info() if method == "fitsne" else None
    t.finish()
    kl_final = t.kl_divergence_pair()
    log(f"full run ({total_epochs} epochs): KL {kl_final}, final grid {final_grid}")

    # the autotuner's epoch form against the built-in one (BHT_FLAG_NO_AUTOTUNE), same epochs
    flags = t.params()["flags"]
    t._set(flags=flags | 4)
    t.initial_embedding(y_seed).prepare(method, 0.5)
    t.synchronize()
    s0 = time.perf_counter()
    t.step(steps)
    t.synchronize()
    builtin_eps = steps / (time.perf_counter() - s0)
    builtin_form = t.epoch_form()
    t.finish()
    t._set(flags=flags)
    autotune = {"epoch_form": form, "builtin_form": builtin_form,
                "builtin_epochs_per_s": round(builtin_eps, 3),
                "gain": round(eps / builtin_eps, 4),
                "meaning": "-1: epoch in order on one stream; k > 0: repulsion beside an "
                           "attraction on k workgroups"}
    log(f"autotuner: form {form} ({eps:.1f} epochs/s) against built-in {builtin_form} "
        f"({builtin_eps:.1f} epochs/s)")

    # the reference's grid (>= 10 boxes) beside the 50-box floor, same timed epochs
    ref_grid = None
    if method == "fitsne" and wl["min_boxes"] != 10:
        t._set(min_boxes=10)
        t.initial_embedding(y_seed).prepare(method, 0.5)
        t.synchronize()
        s0 = time.perf_counter()
        t.step(steps)
        t.synchronize()
        dt = time.perf_counter() - s0
        ref_grid = {"epochs_per_s": round(steps / dt, 3), "min_boxes": 10,
                    "grid": t.grid_info()}
        t.finish()
        t._set(min_boxes=wl["min_boxes"])
        log(f"reference grid (>= 10 boxes): {steps / dt:.1f} epochs/s, {ref_grid['grid']}")

    # dominant kernel: the attraction over the column-tiled P; algorithmic bytes per launch are
    # the edges (packed row/column key + value, 8 B each), y read once and the sums written
    alg_bytes = 8 * nnz + 8 * n + 8 * n
    att_avg_s = att_ms / 1e3 / max(timed, 1)
    achieved = alg_bytes / att_avg_s / 1e9
    forked = n >= 32768 and not args.serial_epoch
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic("k_attract") if args.workload == "c3" else None,
                "algorithmic_bytes": alg_bytes,
                "kernel": "k_attract_tiled<2,true> (attraction sums over the tiled P" +
                          ("; timed while the repulsion runs beside it)" if forked else ")"),
                "avg_ms": round(att_avg_s * 1e3, 4),
                "repulsion_avg_ms": round(rep_ms / max(timed, 1), 4),
                "timing": f"HIP events on the kernels' streams over a second pass of the same "
                          f"{timed} epochs from the seed state"}
    if roofline["traffic"] is None:
        roofline["traffic_note"] = ("no counter pass of this kernel on record; see "
                                    "profiles/k_attract_c3_hbm_traffic.md")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(aff, y_seed, state, late_epoch, wl["min_boxes"], args.cpu_epochs,
                           method)
    del aff, state, t
    batched = serving_line = None
    if args.workload == "c3" and not args.no_batched:
        # the metric's second half: the 256 batched jobs against plain fits of every 4th job,
        # and 32 requests at 24/s through the request loop against plain FIFO fits
        sub = argparse.Namespace(**vars(args))
        sub.jobs, sub.plain_stride, sub.job_epochs = 256, 4, 1000
        b = bench_batched(sub)
        batched = {k: b[k] for k in ("config", "batched", "plain", "vs_plain", "parity")}
        # a saturating arrival rate (the stream's capacity, not its arrival rate, is measured)
        sub.requests, sub.rate, sub.step_epochs = 48, 400.0, 50
        v = bench_serving(sub)
        serving_line = {k: v[k] for k in ("config", "batched", "plain", "vs_plain", "parity")}

    def kl_obj(pair):
        return None if pair is None else {"reference_formula": round(pair[0], 6),
                                          "true": round(pair[1], 6)}

    out = {"metric": METRIC, "value": round(eps, 3), "unit": "epochs/s", "n_gpus": 1,
           "steps": steps, "warmup": args.warmup,
           "ms_per_step": round(1e3 * elapsed / steps, 4), "higher_is_better": True,
           "scaling": "strong", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": wl["name"], "method": method, "n": n,
                      "d_in": int(x.shape[1]), "d_out": 2, "perplexity": 30,
                      "timed_epochs": f"0..{steps} from the seed state (run length "
                                      f"{total_epochs})",
                      "min_boxes": wl["min_boxes"] if method == "fitsne" else None,
                      "nnz": nnz, "affinity_build_s": round(t_aff, 3),
                      "grid_after_timed_epochs": grid, "parallelism": "single GPU",
                      "kl_after_timed_epochs": kl_obj(kl_steps),
                      "kl_full_run": kl_obj(kl_final), "final_grid": final_grid,
                      "reference_grid": ref_grid, "autotune": autotune,
                      "error_vs_fp64_exact": {
                          "state": f"epoch {steps}, 5000 sampled rows",
                          "gradient_rel_l2": float(f"{errs['gradient']:.4e}"),
                          "repulsion_rel_l2": float(f"{errs['repulsion']:.4e}"),
                          "z_rel": float(f"{errs['z']:.4e}")}},
           "roofline": roofline, "cpu_baseline": cpu,
           "batched_jobs": batched, "serving": serving_line}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
