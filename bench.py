"""Benchmark: epochs/s of 2-D FIt-SNE at N = 1,000,000 on one MI355X.

A step is one epoch of the FIt-SNE loop (repulsion by FFT interpolation, CSR attraction fused
with the gains/momentum update, recentring) over the whole embedding.  The workload is the
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
               "grid >= 50 boxes/axis", min_boxes=50),
    "c2": dict(n=70_000, name="FIt-SNE 2D, N=70,000, D_in=20, 10-component GMM (centres "
               "N(0,16I), axis std 1/sqrt(i+1), seed 1), perplexity 30", min_boxes=10),
    "c1": dict(n=5_000, name="FIt-SNE 2D, N=5,000, D_in=50, 10-component GMM (seed 0), "
               "perplexity 30", min_boxes=10),
}


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


def cpu_baseline(aff, y_seed, state_late, min_boxes, n_epochs=3):
    """The oracle (a C restatement of the reference's f32 loop, OpenMP over rows as rayon is)
    timed on this host: n_epochs from the seed state and n_epochs from the GPU run's final state,
    averaged per epoch."""
    import oracle
    threads = oracle.set_threads(0)
    rows = np.ascontiguousarray(aff.rows, np.uint64)
    cols = np.ascontiguousarray(aff.columns, np.uint32)
    n = y_seed.shape[0]
    lr = oracle.auto_learning_rate(n, 12.0)
    times = []
    for label, (y, uy, g), vals_scale in (
            ("early", (y_seed.copy(), np.zeros_like(y_seed), np.ones_like(y_seed)), 12.0),
            ("late", tuple(a.copy() for a in state_late), 1.0)):
        vals = (np.ascontiguousarray(aff.values, np.float32) * np.float32(vals_scale))
        t0 = time.perf_counter()
        oracle.run_epochs(y, uy, g, rows, cols, vals, method="fitsne", min_boxes=min_boxes,
                          learning_rate=lr, e0=0 if label == "early" else 600,
                          e1=n_epochs if label == "early" else 600 + n_epochs)
        dt = time.perf_counter() - t0
        times.append(dt / n_epochs)
        log(f"cpu baseline {label}: {dt / n_epochs:.3f} s/epoch")
    per_epoch = float(np.mean(times))
    return {"value": 1.0 / per_epoch, "unit": "epochs/s", "cores": int(threads), "kind": "port",
            "sample": f"oracle (C restatement of the reference loop, reference grid >= "
                      f"{min_boxes} boxes) on this host: {n_epochs} epochs from the seed state + "
                      f"{n_epochs} epochs from the GPU run's final state, mean s/epoch "
                      f"(early {times[0]:.3f}, late {times[1]:.3f})"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=750)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-epochs", type=int, default=3)
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

    import bhtsne_amd as bt
    from bhtsne_amd import datagen

    wl = WORKLOADS[args.workload]
    t0 = time.perf_counter()
    x = make_data(args.workload)
    n = x.shape[0]
    log(f"data {x.shape} generated in {time.perf_counter() - t0:.1f}s sha256={datagen.sha256(x)[:16]}")
    y_seed = datagen.seed_embedding(n, 2, 2)

    t = bt.TSNE(x).perplexity(30.0).epochs(args.steps).min_boxes(wl["min_boxes"])
    # affinities: kNN + bandwidth search + symmetrisation on the GPU (not in the timed region)
    t0 = time.perf_counter()
    t.initial_embedding(y_seed).prepare("fitsne")
    t.synchronize()
    t_aff = time.perf_counter() - t0
    aff = t.affinities()
    nnz = int(aff.columns.size)
    log(f"affinities built in {t_aff:.2f}s, nnz={nnz}")

    # warmup run (discarded), then a fresh run from the seed state
    if args.warmup > 0:
        t.step(args.warmup)
        t.synchronize()
    t.initial_embedding(y_seed).prepare("fitsne")
    t.set_timing(True)
    t.synchronize()
    start = time.perf_counter()
    t.step(args.steps)
    t.synchronize()
    elapsed = time.perf_counter() - start
    rep_ms, upd_ms, timed = t.timing()
    t.set_timing(False)
    grid = t.grid_info()
    state = t.get_state()
    t.finish()
    eps = args.steps / elapsed
    log(f"{args.steps} epochs in {elapsed:.3f}s -> {eps:.1f} epochs/s; final grid {grid}")

    # dominant kernel: attraction + update, algorithmic bytes per launch
    alg_bytes = 8 * nnz + 4 * (n + 1) + 28 * 2 * n
    upd_avg_s = upd_ms / 1e3 / max(timed, 1)
    achieved = alg_bytes / upd_avg_s / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic("k_attract") if args.workload == "c3" else None,
                "algorithmic_bytes": alg_bytes,
                "kernel": "k_attract<2,true> (attraction + gains/momentum update)",
                "avg_ms": round(upd_avg_s * 1e3, 4),
                "repulsion_avg_ms": round(rep_ms / max(timed, 1), 4)}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(aff, y_seed, state, 10, args.cpu_epochs)

    out = {"metric": METRIC, "value": round(eps, 3), "unit": "epochs/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(1e3 * elapsed / args.steps, 4), "higher_is_better": True,
           "scaling": "strong", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": wl["name"], "n": n, "d_in": int(x.shape[1]), "d_out": 2,
                      "perplexity": 30, "epochs": args.steps, "min_boxes": wl["min_boxes"],
                      "nnz": nnz, "affinity_build_s": round(t_aff, 3),
                      "final_grid": grid, "parallelism": "single GPU"},
           "roofline": roofline, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
