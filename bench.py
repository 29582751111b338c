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


def measured_traffic(log2n, L):
    """HBM bytes per level-split launch from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, each scaled by a 1 GiB copy), for the workload it
    was measured on; None otherwise."""
    path = os.path.join(ROOT, "profiles", "level_split_u32_n268M_L20.md")
    if (log2n, L) != (28, 20) or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch(level_split) = "):
            return float(line.split("=")[1])
    return None


def query_bench(pkg, ctx, dw, nbits, nq):
    """rank_support_v + select_support_mcl (1 and 0) built over all level bits,
    then nq uniform rank(i) and nq uniform select1(j) queries (device batches)."""
    rng = np.random.default_rng(SEED + 1)
    ctx.set_blocking(True)
    for _ in range(2):  # warm
        D, s1, s0 = ctx.rank9_select_build(dw, nbits)
        s1.free(); s0.free(); D.free()
    ctx.timer_start()
    D, s1, s0 = ctx.rank9_select_build(dw, nbits)
    build_ms = ctx.timer_stop()
    di = ctx.to_device(rng.integers(0, nbits + 1, nq, dtype=np.uint64))
    dj = ctx.to_device(rng.integers(1, s1.arg_cnt + 1, nq, dtype=np.uint64))
    out = ctx.empty(nq, np.uint64)
    res = {}
    for name, fn in (("rank", lambda: ctx.rank9_batch(dw, D, nbits, di, out)),
                     ("select1", lambda: ctx.select_batch(s1, dw, dj, out))):
        fn()
        ctx.timer_start()
        for _ in range(3):
            fn()
        t = ctx.timer_stop() / 3
        res[name + "_Mq_s"] = round(nq / (t / 1e3) / 1e6, 1)
    # spot parity on the last select batch: rank(select1(j)+1) == j
    sel = out.get()[:1000]
    dq = ctx.to_device(sel + 1)
    r = ctx.rank9_batch(dw, D, nbits, dq).get()
    ok = bool((r == dj.get()[:1000]).all())
    for a in (di, dj, out, dq):
        a.free()
    s1.free(); s0.free(); D.free()
    res.update({"rank9_select_build_ms": round(build_ms, 3), "nbits": nbits, "queries": nq,
                "rank_select_consistent": ok})
    return res


def bench_lsel(pkg, args):
    """Learned vs plain sampled select: 64 bitvectors of 2^26 bits (densities
    log-uniform in [0.01, 0.5], even Bernoulli, odd runs), uniform (vector, rank)
    queries; both answer sets must be identical."""
    nvec, nb, nq = 64, 1 << args.lsel_log2bits, args.lsel_queries
    log("generating %d bitvectors of %d bits" % (nvec, nb))
    w, dens = pkg.gen_select_vectors(SEED + 5, nvec, nb)
    ctx = pkg.Context(0)
    nw = w.shape[1]
    dv = []
    for b in range(nvec):
        padded = np.zeros(nw + 1, np.uint64)
        padded[:nw] = w[b]
        dv.append(ctx.to_device(padded))
    del w
    log("build + train")
    t0 = time.perf_counter()
    prm = pkg.LselParams(16, args.lsel_gap, args.lsel_tol, args.lsel_epochs, 4.0, 512)
    ls = pkg.LearnedSelect(ctx, dv, [nb] * nvec, prm)
    build_s = time.perf_counter() - t0
    infos = [ls.info(b) for b in range(nvec)]
    rng = np.random.default_rng(SEED + 6)
    vec = rng.integers(0, nvec, nq).astype(np.uint32)
    ones = np.array([i["ones"] for i in infos], np.uint64)
    j = (rng.random(nq) * ones[vec]).astype(np.uint64) + 1
    dvq, djq = ctx.to_device(vec), ctx.to_device(j)
    del vec, j
    out_p, out_l = ctx.empty(nq, np.uint64), ctx.empty(nq, np.uint64)
    res = {}
    for name, learned, out in (("plain", False, out_p), ("learned", True, out_l)):
        ls.query(dvq, djq, out, learned)
        ctx.timer_start()
        for _ in range(args.steps):
            ls.query(dvq, djq, out, learned)
        ms = ctx.timer_stop() / args.steps
        res[name] = {"Mq_s": round(nq / (ms / 1e3) / 1e6, 1), "ms": round(ms, 3)}
    same = bool(np.array_equal(out_p.get(), out_l.get()))
    nl = sum(1 for i in infos if i["learned"])
    model_bytes = sum(i["nknots"] * 8 + ((i["ones"] >> i["lut_log"]) + 1) * 4 for i in infos if i["learned"])
    plain_bytes = sum(i["nbits"] // 8 for i in infos)
    out = {"metric": "select Mq/s learned vs plain (bit-exact)", "value": res["learned"]["Mq_s"],
           "unit": "Mq/s", "n_gpus": 1, "steps": args.steps, "higher_is_better": True,
           "dtype": "u64", "data": "synthetic",
           "config": {"workload": "learned select, 64 bitvectors x 2^%d bits, %d uniform queries"
                                  % (args.lsel_log2bits, nq), "tol": args.lsel_tol, "max_gap": args.lsel_gap,
                      "epochs": args.lsel_epochs},
           "plain_Mq_s": res["plain"]["Mq_s"], "speedup": round(res["learned"]["Mq_s"] / res["plain"]["Mq_s"], 3),
           "identical_answers": same, "learned_vectors": nl,
           "trained_vectors": sum(1 for i in infos if i["learned"] and i["trained"]),
           "model_bytes": model_bytes, "model_bytes_per_bitvector_byte": round(model_bytes / plain_bytes, 4),
           "build_and_train_s": round(build_s, 3),
           "eps_max_learned": max([i["eps"] for i in infos if i["learned"]] or [0]),
           "per_vector": [{"d": round(float(dens[b]), 4), "learned": infos[b]["learned"],
                           "trained": infos[b]["trained"], "nknots": infos[b]["nknots"],
                           "eps": infos[b]["eps"], "eps_init": infos[b]["eps_init"],
                           "loss_init": round(infos[b]["loss_init"], 1),
                           "loss_final": round(infos[b]["loss_final"], 1)} for b in range(nvec)]}
    ls.free()
    for a in dv + [dvq, djq, out_p, out_l]:
        a.free()
    ctx.close()
    print(json.dumps(out), flush=True)
    return 0 if same else 1


def bench_batch(pkg, args):
    """Batched level-wise trees: 4096 byte sequences, n_i = floor(2^U(10,22)),
    alphabets {4, 20, 256}; one hipGraph replay per step, against per-sequence
    plain launches and the reference levelWT on a bounded sample."""
    import ctypes as C
    nseq = args.batch_seqs
    log("generating %d sequences" % nseq)
    n, sigma, sym = pkg.gen_batch(SEED + 7, nseq, 10.0, float(args.batch_log2max))
    total = int(n.sum())
    off = np.concatenate([[0], np.cumsum(n)[:-1]]).astype(np.uint64)
    ctx = pkg.Context(0)
    d = ctx.to_device(sym)
    # the autotuner picks each alphabet width's group size on these sequences
    t0 = time.perf_counter()
    plan, tune_ms = pkg.Batch.tune(ctx, d, n, sigma, off, reps=2)
    tune_s = time.perf_counter() - t0
    B = pkg.Batch(ctx, d, n, sigma, off, plan=plan)
    L = np.array([pkg.levels_of_sigma(int(x)) for x in sigma])
    log("batch: %d symbols, plan %s" % (total, list(plan)))
    B.run()
    ctx.set_blocking(False)
    for _ in range(args.warmup):
        B.run()
    ctx.sync()
    ctx.timer_start()
    for _ in range(args.steps):
        B.run()
    ms = ctx.timer_stop() / args.steps
    ctx.sync()
    # per-sequence plain launches into one output array
    wo = np.concatenate([[0], np.cumsum([pkg.lib().wtg_levels_words(int(k), int(l)) + 1
                                         for k, l in zip(n, L)])]).astype(np.uint64)
    dw = ctx.empty(int(wo[-1]), np.uint64)
    lib = pkg.lib()

    def per_seq():
        for i in range(nseq):
            if n[i]:
                pkg._check(lib.wtg_levels_build(ctx.h, C.c_void_p(d.ptr.value + int(off[i])), 1, int(n[i]),
                                                int(L[i]), C.c_void_p(dw.ptr.value + 8 * int(wo[i])), None, 0))
    per_seq()
    ctx.sync()
    t0 = time.perf_counter()
    ctx.timer_start()
    per_seq()
    ms_seq = ctx.timer_stop()
    ctx.sync()
    wall_seq = time.perf_counter() - t0
    ctx.set_blocking(True)
    # parity: the batch's sequences equal the per-sequence builds
    allw = dw.get()
    ok = True
    rng = np.random.default_rng(1)
    for i in list(rng.choice(nseq, 24, replace=False)) + [int(np.argmax(n))]:
        e = allw[int(wo[i]):int(wo[i]) + pkg.lib().wtg_levels_words(int(n[i]), int(L[i]))]
        ok &= bool(np.array_equal(B.extract(int(i)), e))
    del allw
    B_a = float((n.astype(np.float64) * (2 * L + L / 8.0)).sum())
    out = {"metric": "batched WT build Gsym/s", "value": round(total / (ms / 1e3) / 1e9, 4), "unit": "Gsym/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
           "higher_is_better": True, "dtype": "u8", "data": "synthetic",
           "config": {"workload": "batched levelWT, %d sequences, n_i=floor(2^U(10,%d)), alphabets {4,20,256}"
                                  % (nseq, args.batch_log2max), "symbols": total},
           "build_roofline": {"B_a_bytes": B_a, "frac": round(B_a / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)},
           "per_sequence_launches": {"ms": round(ms_seq, 3), "wall_ms": round(wall_seq * 1e3, 3),
                                     "Gsym_s": round(total / (ms_seq / 1e3) / 1e9, 4)},
           "speedup_vs_per_sequence": round(ms_seq / ms, 2), "parity_vs_per_sequence": ok,
           "autotune": {"cap_log_by_width": {int(w): int(plan[w]) for w in sorted(set(L.tolist()))},
                        "candidates_ms": {int(w): [round(float(x), 3) for x in tune_ms[w]]
                                          for w in sorted(set(L.tolist()))},
                        "tune_s": round(tune_s, 3),
                        "default_16bit_keys": {int(w): 16 - int(w) for w in sorted(set(L.tolist()))}}}
    if not args.no_cpu:
        from oracle import oracle
        ref = oracle.RefWT("levelWT", "u8")
        acc_n, acc_t = 0, 0.0
        for i in range(nseq):
            if acc_n >= (1 << 26):
                break
            if n[i] == 0:
                continue
            s_i = sym[int(off[i]):int(off[i]) + int(n[i])]
            if ref.ok:
                acc_t += float(ref.time(s_i, 1 << int(L[i]), rounds=1).min())
            else:
                t0 = time.perf_counter()
                oracle.levels_serial(s_i, int(L[i]))
                acc_t += time.perf_counter() - t0
            acc_n += int(n[i])
        out["cpu_baseline"] = {"value": round(acc_n / acc_t / 1e9, 5), "unit": "Gsym/s",
                               "cores": int(ref.workers()) if ref.ok else 1,
                               "kind": "reference" if ref.ok else "port",
                               "sample": "the first %d sequences' symbols (%d), one WT() each" % (i, acc_n)}
    B.free(); d.free(); dw.free(); ctx.close()
    print(json.dumps(out), flush=True)
    return 0 if ok else 1


def bench_fm(pkg, args):
    """FM-index over synthetic DNA: host SA + BWT, wt_huff build on the GPU (timed
    per build, input resident), then batched backward-search counts of
    length-20 patterns (half exact substrings, half with 1-2 substitutions)."""
    n = 1 << args.fm_log2n
    log("generating %d DNA bytes" % n)
    t = pkg.gen_dna(SEED + 3, n)
    t0 = time.perf_counter()
    sa = pkg.suffix_array(t)
    sa_s = time.perf_counter() - t0
    bwt = pkg.bwt_from_sa(t, sa)
    pats = pkg.gen_patterns(SEED + 4, t, 20, args.fm_patterns)
    ctx = pkg.Context(0)
    # the same suffix array on the GPU (prefix doubling), checked against the host's
    dt = ctx.to_device(t)
    pkg.suffix_array_gpu(ctx, dt, n).free()
    ctx.sync()
    t0 = time.perf_counter()
    dsa = pkg.suffix_array_gpu(ctx, dt, n)
    ctx.sync()
    gpu_sa_s = time.perf_counter() - t0
    sa_match = bool(np.array_equal(dsa.get(), sa))
    dsa.free(); dt.free()
    d = ctx.to_device(bwt)
    for _ in range(max(args.warmup, 1)):
        wt = ctx.huff_build(d)
        wt.free()
    ctx.timer_start()
    for _ in range(args.steps):
        wt = ctx.huff_build(d)
        wt.free()
    build_ms = ctx.timer_stop() / args.steps
    # per-kernel breakdown in separate builds (the per-launch events cost host time)
    ctx.kernel_timing_reset()
    ctx.kernel_timing(True)
    for _ in range(args.steps):
        wt = ctx.huff_build(d)
        wt.free()
    kt = {k: ctx.kernel_time(k) for k in ("huff_hist", "huff_count", "huff_scan", "huff_emit", "rank_select")}
    ctx.kernel_timing(False)
    wt = ctx.huff_build(d)
    dp = ctx.to_device(pats.reshape(-1))
    do = ctx.empty(pats.shape[0], np.uint64)
    lib = pkg.lib()
    pkg._check(lib.wtg_fm_count_batch(ctx.h, wt.h, dp.ptr, 20, pats.shape[0], do.ptr))
    ctx.timer_start()
    for _ in range(args.steps):
        pkg._check(lib.wtg_fm_count_batch(ctx.h, wt.h, dp.ptr, 20, pats.shape[0], do.ptr))
    q_ms = ctx.timer_stop() / args.steps
    got = do.get()
    from oracle import oracle
    ok = bool(np.array_equal(got, oracle.sa_count(t, sa, pats)))
    bits = int(wt.shape.bv_size)
    B_c = n * 1.0 + 2.0 * bits + bits / 8.0   # n*w + sum_l a_l*2w + n*H/8 (w = 1, sum_l a_l = bits)
    emit_ms = kt["huff_emit"][0] / max(kt["huff_emit"][1], 1)
    out = {"metric": "wt_huff build over a BWT Gsym/s; FM count Mpat/s", "value": round(n / (build_ms / 1e3) / 1e9, 4),
           "unit": "Gsym/s", "n_gpus": 1, "steps": args.steps, "higher_is_better": True, "dtype": "u8",
           "data": "synthetic", "config": {"workload": "FM-index over 2^%d-byte order-5 Markov DNA, %d length-20 patterns"
                                                      % (args.fm_log2n, pats.shape[0])},
           "build_ms": round(build_ms, 4), "bits": bits, "mean_code_len": round(bits / n, 4),
           "build_roofline": {"B_c_bytes": B_c, "frac": round(B_c / (build_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)},
           "kernels_ms": {k: round(v[0] / max(v[1], 1), 4) for k, v in kt.items()},
           "emit_GBps": round((n + bits / 8.0) / (emit_ms / 1e3) / 1e9, 1) if emit_ms else None,
           "count_ms": round(q_ms, 4), "count_Mpat_s": round(pats.shape[0] / (q_ms / 1e3) / 1e6, 2),
           "counts_match_sa_binary_search": ok, "host_sa_s": round(sa_s, 2), "gpu_sa_s": round(gpu_sa_s, 4),
           "gpu_sa_matches_host": sa_match}
    if not args.no_cpu:
        # CPU: the oracle's serial restatement of the Huffman tree bits (wt_huff_par.hpp:92-117,
        # wt_pc_par.hpp:193-257) on a prefix of the BWT
        m = min(n, 1 << 25)
        t0 = time.perf_counter()
        oracle.Huff(bwt[:m])
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(m / dt / 1e9, 5), "unit": "Gsym/s", "cores": 1, "kind": "port",
                               "sample": "serial Huffman-shape + tree bits + rank9 + select restatement on the "
                                         "first %d BWT bytes" % m}
    wt.free(); d.free(); dp.free(); do.free(); ctx.close()
    print(json.dumps(out), flush=True)
    return 0 if ok and sa_match else 1


def bench_bytes(pkg, args):
    """Byte wavelet tree (levelWT): n = 2^24 uint8 Zipf(1.1) symbols over a seeded
    byte permutation, 8 levels + rank directory per step; 10^6 uniform rank(c, i)
    queries; bit-exact against the compiled reference serialWT."""
    n, L = 1 << 24, 8
    s = pkg.gen_zipf_bytes(SEED + 1, n, 1.1)
    ctx = pkg.Context(0)
    d = ctx.to_device(s)
    nw = pkg.lib().wtg_levels_words(n, L)
    dw = ctx.empty(nw, np.uint64)
    D = ctx.empty(pkg.lib().wtg_rank9_words(n * L), np.uint64)
    lib = pkg.lib()

    def step():
        ctx.levels_build(d, n, L, dw)
        pkg._check(lib.wtg_rank9_build(ctx.h, dw.ptr, n * L, D.ptr))
    ctx.set_blocking(False)
    for _ in range(args.warmup):
        step()
    ctx.sync()
    ctx.timer_start()
    for _ in range(args.steps):
        step()
    ms = ctx.timer_stop() / args.steps
    ctx.sync()
    ctx.set_blocking(True)
    words = dw.get()
    from oracle import oracle
    ref = oracle.RefWT("serialWT", "u8")
    exp = ref.build(s, 256)[0] if ref.ok else oracle.levels_serial(s, L)[0]
    ok = bool(np.array_equal(words, exp))
    wt = ctx.wtint_build(d, n, L)
    rng = np.random.default_rng(SEED + 2)
    qi = rng.integers(0, n + 1, 1_000_000).astype(np.uint64)
    qc = rng.integers(0, 256, 1_000_000).astype(np.uint64)
    di, dc, do = ctx.to_device(qi), ctx.to_device(qc), ctx.empty(qi.size, np.uint64)
    pkg._check(lib.wtg_wtint_rank_batch(ctx.h, wt.h, di.ptr, dc.ptr, qi.size, do.ptr))
    ctx.timer_start()
    for _ in range(args.steps):
        pkg._check(lib.wtg_wtint_rank_batch(ctx.h, wt.h, di.ptr, dc.ptr, qi.size, do.ptr))
    q_ms = ctx.timer_stop() / args.steps
    got = do.get()
    for k in range(0, 1000):
        c, i = int(qc[k]), int(qi[k])
        ok &= int(got[k]) == int((s[:i] == c).sum())
    B_a = n * (2 * L * 1 + L / 8.0)
    out = {"metric": "byte WT build Gsym/s", "value": round(n / (ms / 1e3) / 1e9, 4), "unit": "Gsym/s",
           "n_gpus": 1, "steps": args.steps, "higher_is_better": True, "dtype": "u8", "data": "synthetic",
           "config": {"workload": "levelWT bytes, n=2^24 Zipf(1.1) over a byte permutation, 8 levels + rank9"},
           "ms_per_step": round(ms, 4),
           "build_roofline": {"B_a_bytes": B_a, "frac": round(B_a / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)},
           "rank_queries_Mq_s": round(qi.size / (q_ms / 1e3) / 1e6, 1),
           "bit_exact_vs_reference_serialWT": ok}
    if not args.no_cpu:
        lw = oracle.RefWT("levelWT", "u8")
        if lw.ok:
            t = float(lw.time(s, 256, rounds=3).min())
            out["cpu_baseline"] = {"value": round(n / t / 1e9, 5), "unit": "Gsym/s", "cores": int(lw.workers()),
                                   "kind": "reference", "sample": "reference levelWT (OpenMP, byte build) on the "
                                                                  "whole input, best of 3 (tree only)"}
    for a in (d, dw, D, di, dc, do):
        a.free()
    wt.free()
    ctx.close()
    print(json.dumps(out), flush=True)
    return 0 if ok else 1


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log2n", type=int, default=28)
    ap.add_argument("--levels", type=int, default=20)
    ap.add_argument("--cpu-log2n", type=int, default=26)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first copy 1 GiB device-to-device twice (a plain copy of known size, "
                         "the reference point for HBM byte counters)")
    ap.add_argument("--no-geometric", dest="geometric", action="store_false",
                    help="skip the geometric(1e-4) second input")
    ap.add_argument("--queries", type=int, default=10_000_000,
                    help="rank and select queries on the built tree (0: skip)")
    ap.add_argument("--workload", default="wt_int", choices=["wt_int", "bytes", "lsel", "batch", "fm"],
                    help="wt_int: the tree build (the headline line); lsel: learned vs plain select")
    ap.add_argument("--lsel-log2bits", type=int, default=26)
    ap.add_argument("--batch-seqs", type=int, default=4096)
    ap.add_argument("--fm-log2n", type=int, default=27)
    ap.add_argument("--fm-patterns", type=int, default=1_000_000)
    ap.add_argument("--batch-log2max", type=int, default=22)
    ap.add_argument("--lsel-queries", type=int, default=100_000_000)
    ap.add_argument("--lsel-tol", type=int, default=480, help="learned select: knot placement tolerance (bits)")
    ap.add_argument("--lsel-gap", type=int, default=4096, help="learned select: longest knot interval (ones)")
    ap.add_argument("--lsel-epochs", type=int, default=200, help="learned select: Adam steps")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the level build of this path runs on a single "
              "MI355X; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        return 2
    pkg = load_pkg()
    if args.workload == "lsel":
        return bench_lsel(pkg, args)
    if args.workload == "bytes":
        return bench_bytes(pkg, args)
    if args.workload == "batch":
        return bench_batch(pkg, args)
    if args.workload == "fm":
        return bench_fm(pkg, args)
    n, L, w = 1 << args.log2n, args.levels, 4
    log("generating n=%d uint32 symbols in [0,2^%d)" % (n, L))
    s = pkg.gen_uniform(SEED, n, 1 << L, np.uint32)
    ctx = pkg.Context(0)
    if args.calibrate_copy:
        a, b = ctx.empty(1 << 27, np.uint64), ctx.empty(1 << 27, np.uint64)
        a.zero()
        for _ in range(2):
            pkg._check(pkg.lib().wtg_memcpy_d2d(ctx.h, b.ptr, a.ptr, 1 << 30))
        ctx.sync()
        a.free(); b.free()
    d = ctx.to_device(s)
    nw = pkg.lib().wtg_levels_words(n, L)
    dw = ctx.empty(nw, np.uint64)
    ctx.set_blocking(False)
    log("warmup %d" % args.warmup)
    for _ in range(args.warmup):
        ctx.levels_build(d, n, L, dw)
    ctx.sync()
    ctx.kernel_timing_reset()
    ctx.kernel_timing_only("level_")  # events around the level kernels only
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

    extras = query_bench(pkg, ctx, dw, n * L, args.queries) if args.queries else None

    # size-independent parity check of the timed output: ones per level
    words = dw.get()
    bits = np.unpackbits(words.view(np.uint8), bitorder="little")
    ok = True
    for l in range(L):
        lv = bits[l * n:(l + 1) * n]
        if int(lv.sum()) != int(((s >> (L - 1 - l)) & 1).sum()):
            ok = False
    del bits, words
    # the second input of the same workload: geometric(p = 1e-4) values capped at
    # 2^20 - 1, the same 20 levels (wt_int max_level = 20)
    geo = None
    if args.geometric:
        g = pkg.gen_geometric(SEED + 9, n, 1e-4, (1 << L) - 1, np.uint32)
        dg = ctx.to_device(g)
        ctx.levels_build(dg, n, L, dw)
        ctx.set_blocking(False)
        ctx.timer_start()
        for _ in range(3):
            ctx.levels_build(dg, n, L, dw)
        gms = ctx.timer_stop() / 3
        ctx.sync()
        ctx.set_blocking(True)
        gw = dw.get()
        gok = True
        gb = np.unpackbits(gw.view(np.uint8), bitorder="little")
        for l in range(L):
            if int(gb[l * n:(l + 1) * n].sum()) != int(((g >> (L - 1 - l)) & 1).sum()):
                gok = False
        del gb, gw
        dg.free()
        geo = {"Gsym_s": round(n / (gms / 1e3) / 1e9, 4), "ms": round(gms, 4), "level_popcounts_match": gok}

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
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(args.log2n, L),
                     "kernel": "k_split<uint32> (level_split)",
                     "bytes_per_launch": split_bytes, "avg_launch_ms": round(split_avg_s * 1e3, 4),
                     "launches": split_n},
        "build_roofline": {"B_a_bytes": B_a, "frac": round(B_a / (step_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                           "level_bits_avg_ms": round(bits_ms / max(bits_n, 1), 4),
                           "hist_ms_per_step": round(hist_ms / args.steps, 4)},
        "parity": {"level_popcounts_match": ok},
        "rank_select": extras,
        "geometric_input": geo,
        "wall_s": round(wall, 3),
    }
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(s, L, 1 << args.cpu_log2n, 3)
    print(json.dumps(out), flush=True)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
