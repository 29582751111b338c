#!/usr/bin/env python3
"""Benchmark of the level-wise wavelet-tree build on one MI355X.

Workload (default): the integer wavelet tree of sdsl::wt_int<> with
max_level = 20 over n = 2^28 uint32 symbols drawn uniformly from [0, 2^20)
(seed 0x5EED2026).  One step = one full tree build (histogram, node tables and
the 20 level passes) with the input already resident in HBM.

Prints ONE JSON line: metric/value (Gsym/s = n / step time), the roofline of
the dominant kernel (the level split pass, algorithmic bytes 2*n*w + n/8 per
launch over its HIP-event-timed average duration), the whole-build fraction of
HBM peak against B_a = n*(2*L*w + L/8), and a CPU baseline: the reference
levelWT (wavelet/levelWT/WT.C, compiled with its OpenMP backend and
-DINT -DLONG into oracle/_ref) timed on this host's cores on a bounded prefix
of the same input.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "WT build Gsym/s & %HBM peak; select Mq/s learned vs plain (bit-exact)"
SEED = 0x5EED2026


def load_pkg():
    import importlib.util
    d = os.path.join(ROOT, "wavelet-suffix-fm-index_amd")
    name = "wavelet_suffix_fm_index_amd"
    spec = importlib.util.spec_from_file_location(name, os.path.join(d, "__init__.py"),
                                                  submodule_search_locations=[d])
    m = importlib.util.module_from_spec(spec)
    sys.modules[name] = m
    spec.loader.exec_module(m)
    return m


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(s, L, sample_n, rounds):
    """Reference levelWT (oracle/_ref, OpenMP) on the first sample_n symbols."""
    from oracle import oracle
    sample = np.ascontiguousarray(s[:sample_n].astype(np.uint64))
    ref = oracle.RefWT("levelWT", "u64")
    if ref.ok:
        t = ref.time(sample, 1 << L, rounds=rounds)
        kind, cores = "reference", ref.workers()
        best = float(t.min())
    else:  # the oracle's serial restatement (1 core)
        t0 = time.perf_counter()
        oracle.levels_serial(sample, L, with_nodes=False)
        best = time.perf_counter() - t0
        kind, cores = "port", 1
    return {"value": round(sample_n / best / 1e9, 5), "unit": "Gsym/s", "cores": int(cores),
            "kind": kind,
            "sample": "reference levelWT (OpenMP, -DINT -DLONG) on the first %d symbols of the "
                      "same input, L=%d, best of %d rounds" % (sample_n, L, rounds)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log2n", type=int, default=28)
    ap.add_argument("--levels", type=int, default=20)
    ap.add_argument("--cpu-log2n", type=int, default=26)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the level build of this path runs on a single "
              "MI355X; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        return 2
    pkg = load_pkg()
    n, L, w = 1 << args.log2n, args.levels, 4
    log("generating n=%d uint32 symbols in [0,2^%d)" % (n, L))
    s = pkg.gen_uniform(SEED, n, 1 << L, np.uint32)
    ctx = pkg.Context(0)
    d = ctx.to_device(s)
    nw = pkg.lib().wtg_levels_words(n, L)
    dw = ctx.empty(nw, np.uint64)
    ctx.set_blocking(False)
    log("warmup %d" % args.warmup)
    for _ in range(args.warmup):
        ctx.levels_build(d, n, L, dw)
    ctx.sync()
    ctx.kernel_timing_reset()
    ctx.kernel_timing(True)
    ctx.sync()
    log("timing %d steps" % args.steps)
    t0 = time.perf_counter()
    ctx.timer_start()
    for _ in range(args.steps):
        ctx.levels_build(d, n, L, dw)
    ms = ctx.timer_stop()
    ctx.sync()
    wall = time.perf_counter() - t0
    split_ms, split_n = ctx.kernel_time("level_split")
    bits_ms, bits_n = ctx.kernel_time("level_bits")
    hist_ms, _ = ctx.kernel_time("hist")
    ctx.kernel_timing(False)
    ctx.set_blocking(True)

    # size-independent parity check of the timed output: ones per level
    words = dw.get()
    bits = np.unpackbits(words.view(np.uint8), bitorder="little")
    ok = True
    for l in range(L):
        lv = bits[l * n:(l + 1) * n]
        if int(lv.sum()) != int(((s >> (L - 1 - l)) & 1).sum()):
            ok = False
    del bits, words
    d.free()
    dw.free()
    ctx.close()

    step_ms = ms / args.steps
    value = n / (step_ms / 1e3) / 1e9
    B_a = n * (2 * L * w + L / 8.0)
    split_bytes = 2.0 * n * w + n / 8.0
    split_avg_s = split_ms / max(split_n, 1) / 1e3
    achieved = split_bytes / split_avg_s / 1e9
    out = {
        "metric": METRIC, "value": round(value, 4), "unit": "Gsym/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(step_ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "wt_int tree build (levelWT level split), n=2^%d uint32 uniform "
                               "in [0,2^%d), max_level=%d" % (args.log2n, L, L),
                   "n": n, "levels": L, "sym_bytes": w, "seed": SEED},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_level<uint32,scatter> (level_split)",
                     "bytes_per_launch": split_bytes, "avg_launch_ms": round(split_avg_s * 1e3, 4),
                     "launches": split_n},
        "build_roofline": {"B_a_bytes": B_a, "frac": round(B_a / (step_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                           "level_bits_avg_ms": round(bits_ms / max(bits_n, 1), 4),
                           "hist_ms_per_step": round(hist_ms / args.steps, 4)},
        "parity": {"level_popcounts_match": ok},
        "wall_s": round(wall, 3),
    }
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(s, L, 1 << args.cpu_log2n, 3)
    print(json.dumps(out), flush=True)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
