"""Benchmark: wavelet matrix / tree construction throughput on one MI355X.

Default workload (the headline configuration): a 4 GiB uint8 DNA-like text
(codes 0..4, order-2 Markov source, N runs), 3 levels, wavelet matrix built
with the prefix-counting algorithm's drop-in (wm_ppc).  A step is one full
construction of all levels from the text resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config NAME]

Prints one JSON line.  `value` is input GB/s (n * bytes per symbol / time per
step); `roofline` is that of the dominant kernel, the local build (it reads
the text once and writes every level's bits once: n*width + levels*n/8
bytes), over its own device time (HIP events on the build stream);
`config.build_roofline` relates the same algorithmic bytes to the device
time of the whole build; `verified` says whether the last build equals the
reference's own output on the same text (per-level SHA-256 digests, zeros,
codes; tests/golden/fullsize_digests.json), checked after the timed region;
on the headline workload `batched` holds the many-texts comparison (20,000
uneven DNA texts: one batched build against one plain build per text, the
checked structures against per-text builds and the oracle); `cpu_baseline`
is the reference's own OpenMP algorithms (compiled from its headers,
oracle/_ref) on a bounded prefix of the same text, on this host's cores.
The process exits with status 1 if a verification fails.
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

CONFIGS = {
    # name: (generator kind, seed, n, width, levels, algo, tree)
    "dna4g_wm": ("dna", 0xC2, 1 << 32, 1, 3, "ppc", False),
    "dna4g_wt": ("dna", 0xC2, 1 << 32, 1, 3, "ppc", True),
    "uniform64m_wm": ("uniform", 0xC1, 1 << 26, 1, 8, "ppc", False),
    "uniform64m_wt": ("uniform", 0xC1, 1 << 26, 1, 8, "pps", True),
    "zipf1g_wm": ("zipf", 0xC3, 1 << 28, 4, 20, "pps", False),
    "dna4g_huffwm": ("dna", 0xC2, 1 << 32, 1, 0, "huff_pc", False),
}
# the same workloads with the text in pinned host memory, streamed to the
# device in slices during the build (PCIe-bound)
STREAM_CONFIGS = {"dna4g_stream": "dna4g_wm"}
# random symbol queries timed after the build
QUERIES = {"zipf1g_wm": 10**7}
# workloads whose line also carries the many-texts figures
BATCH_WITH = {"dna4g_wm": "dna_batch20k"}


# Many uneven texts, built together: (seed, texts, median length, sigma,
# min length, max length, levels, algo, tree)
BATCH_CONFIGS = {
    "dna_batch20k": (0xC4, 20000, 65536, 1.5, 1024, 64 << 20, 3, "ppc", False),
}


def _omp_env(cores):
    """OpenMP placement for the reference's CPU algorithms: one thread per
    core, spread over the sockets (read by libgomp when the compiled
    reference is first loaded)."""
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    os.environ.setdefault("OMP_PROC_BIND", "spread")
    os.environ.setdefault("OMP_PLACES", "cores")


def _timed_runs(fn, runs=5):
    """One warm-up call, then `runs` timed calls: (true median, minimum)."""
    fn()
    ts = sorted(fn() for _ in range(runs))
    return ts[len(ts) // 2], ts[0]


def cpu_baseline(text, levels, tree, width, sample_bytes, algo="ppc"):
    """Reference algorithms (wx_ppc / wx_pps / wx_dd_pc, or wx_huff_pc for a
    Huffman-shaped build) on a prefix sample: one warm-up and five timed
    runs each, true median (the reference's own "median" is the middle run
    by start time, stats.hpp:162-164)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    cores = len(os.sched_getaffinity(0))
    _omp_env(cores)
    import oracle_lib
    sample = np.ascontiguousarray(text[: sample_bytes // width])
    if oracle_lib.have_ref() and algo.startswith("huff"):
        med, mn = _timed_runs(lambda: oracle_lib.ref_build("huff_pc", sample, 0, tree)[1])
        return {"value": round(sample.nbytes / med / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "reference",
                "sample": "%d MiB prefix of the same text; wx_huff_pc (sequential in the reference), "
                          "1 warm-up + 5 runs, median (min %.3f s)" % (sample.nbytes >> 20, mn)}
    if oracle_lib.have_ref():
        best = None
        per = {}
        for a in ("ppc", "pps", "dd_pc"):
            med, mn = _timed_runs(lambda: oracle_lib.ref_build(a, sample, levels, tree)[1])
            per[a] = {"median_GBps": round(sample.nbytes / med / 1e9, 3), "best_GBps": round(sample.nbytes / mn / 1e9, 3)}
            if best is None or per[a]["median_GBps"] > best[1]:
                best = (a, per[a]["median_GBps"])
        return {"value": best[1], "unit": "GB/s", "cores": int(os.environ["OMP_NUM_THREADS"]),
                "kind": "reference",
                "sample": "%d MiB prefix of the same text; wx_%s, best of ppc/pps/dd_pc by median of 5 runs "
                          "after 1 warm-up, OMP_PROC_BIND=%s OMP_PLACES=%s: %s"
                          % (sample.nbytes >> 20, best[0], os.environ["OMP_PROC_BIND"], os.environ["OMP_PLACES"],
                             per)}
    small = sample[: 16 << 20]
    t0 = time.perf_counter()
    if algo.startswith("huff"):
        oracle_lib.huff_pc(small, tree)
    else:
        oracle_lib.naive(small, levels, tree)
    dt = time.perf_counter() - t0
    return {"value": round(small.nbytes / dt / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": "%d MiB prefix, scalar CPU restatement (reference build absent)" % (small.nbytes >> 20)}


def traffic_from_profiles(config):
    """HBM bytes per local-build launch from the committed counter summary."""
    p = os.path.join(ROOT, "profiles", "%s_traffic.json" % config)
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f).get("local_build_hbm_bytes")
    return None


def rooflines(b_min, phases_ms, total_ms, config, passes=1):
    """Dominant kernel (local build, phase 0) and whole-build figures.  With
    one pass the local build reads the text once and writes every level
    once, so its algorithmic bytes are the build's B_min.  A multi-pass
    build's phase-0 time is its first pass only and does a fraction of B_min;
    there the whole build is reported against B_min instead."""
    build = b_min / (total_ms / 1e3) / 1e9
    whole = {"achieved": round(build, 1), "frac": round(build / HBM_PEAK_GBS, 4), "device_ms": round(total_ms, 4)}
    if passes == 1:
        k1 = b_min / (phases_ms[0] / 1e3) / 1e9
        roof = {"bound": "hbm", "kernel": "local build (k_local_direct / k_local_super / k_local_small)", "achieved": round(k1, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(k1 / HBM_PEAK_GBS, 4),
                "traffic": traffic_from_profiles(config), "algorithmic_bytes": int(b_min),
                "kernel_ms": round(phases_ms[0], 4)}
    else:
        roof = {"bound": "hbm", "kernel": "whole build (%d passes; no single launch does the build's bytes)" % passes,
                "achieved": whole["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": whole["frac"],
                "traffic": None, "algorithmic_bytes": int(b_min), "kernel_ms": round(total_ms, 4)}
    return roof, whole


# Full-size digests of the reference's own outputs (tests/golden/
# fullsize_digests.json): workload -> (text entry, build entry)
DIGESTS = {
    "dna4g_wm": ("dna4g", "wm_ppc"), "dna4g_wt": ("dna4g", "wt_ppc"), "dna4g_huffwm": ("dna4g", "wm_huff_pc"),
    "zipf1g_wm": ("zipf256m", "wm_pps"), "uniform64m_wm": ("uniform64m", "wm_ppc"),
    "uniform64m_wt": ("uniform64m", "wt_pps"), "dna4g_stream": ("dna4g", "wm_ppc"),
}


def verify_against_reference(config, text, ws, lib, ctx):
    """Compares the last build with the reference's own output on the same
    text: per-level SHA-256 of the words, level sizes, zeros and (Huffman)
    codes.  Runs outside the timed region."""
    import ctypes
    import hashlib
    if config not in DIGESTS:
        return {"verified": None, "why": "no reference digest for this workload"}
    with open(os.path.join(ROOT, "tests", "golden", "fullsize_digests.json")) as f:
        dig = json.load(f)
    tname, bname = DIGESTS[config]
    ent = dig[tname]
    if text.size != ent["symbols"]:
        return {"verified": None, "why": "text length differs from the digested workload"}

    def sha(a):
        return hashlib.sha256(memoryview(np.ascontiguousarray(a)).cast("B")).hexdigest()

    rec = ent["builds"][bname]
    bad = []
    if sha(text) != ent["text_sha256"]:
        bad.append("text")
    L = ws.levels()
    if L != len(rec["level_bits"]) or [ws.level_bit_size(l) for l in range(L)] != rec["level_bits"]:
        bad.append("level sizes")
    else:
        for l in range(L):
            a = np.empty((ws.level_bit_size(l) + 63) // 64, np.uint64)
            st = lib.pwmhip_ws_download(ctx.h, ctypes.byref(ws._ws), l, a.ctypes.data_as(ctypes.c_void_p))
            if st or sha(a) != rec["level_sha256"][l]:
                bad.append("level %d" % l)
    if [int(z) for z in ws.zeros()] != rec["zeros"]:
        bad.append("zeros")
    if "code_len" in rec:
        ln, wd = ws.codes()
        if [int(x) for x in ln] != rec["code_len"] or [int(x) for x in wd] != rec["code_word"]:
            bad.append("codes")
    out = {"verified": not bad,
           "against": "reference wx_%s on the same text: per-level SHA-256, level sizes, zeros%s "
                      "(tests/golden/fullsize_digests.json)" % (bname[3:], ", codes" if "code_len" in rec else "")}
    if bad:
        out["mismatch"] = bad
    return out


def batch_sizes(seed, T, median, sigma, lo, hi):
    rng = np.random.default_rng(seed)
    return np.clip(rng.lognormal(np.log(median), sigma, T), lo, hi).astype(np.int64)


def cpu_baseline_batch(host, offs, sizes, levels, tree, sample_bytes):
    """Reference algorithms text by text (one call per text, as the reference
    builds) over the first texts of the batch up to `sample_bytes`."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    if not oracle_lib.have_ref():
        return None
    cores = len(os.sched_getaffinity(0))
    _omp_env(cores)
    k = int(np.searchsorted(np.cumsum(sizes), sample_bytes)) + 1
    k = min(k, len(sizes))
    sample = [host[offs[i]:offs[i] + sizes[i]] for i in range(k)]
    nbytes = int(sum(x.size for x in sample))
    per, best = {}, None
    for algo in ("ppc", "pps", "dd_pc"):
        oracle_lib.ref_build(algo, sample[0], levels, tree)
        t = sum(oracle_lib.ref_build(algo, x, levels, tree)[1] for x in sample)
        per[algo] = round(nbytes / t / 1e9, 3)
        if best is None or per[algo] > best[1]:
            best = (algo, per[algo])
    return {"value": best[1], "unit": "GB/s", "cores": int(os.environ["OMP_NUM_THREADS"]), "kind": "reference",
            "sample": "first %d texts of the batch (%d MiB), one reference call per text; wx_%s best of "
                      "ppc/pps/dd_pc: %s" % (k, nbytes >> 20, best[0], per)}


def batch_figures(config, steps, warmup, graph=True, per_text=True, check_oracle=True):
    """Batched construction of many uneven texts against one build per text:
    returns (figures dict, b_min, phases, device ms, host texts)."""
    import ctypes
    import torch
    import pwm_amd

    seed, T, median, sigma, lo, hi, levels, algo, tree = BATCH_CONFIGS[config]
    sizes = batch_sizes(seed, T, median, sigma, lo, hi)
    padded = (sizes + 255) // 256 * 256
    offs = np.concatenate([[0], np.cumsum(padded)[:-1]]).astype(np.int64)
    total = int(sizes.sum())
    t0 = time.perf_counter()
    host = pwm_amd.generate("dna", seed, int(padded.sum()))  # texts sit at 256-byte aligned offsets
    gen_s = time.perf_counter() - t0
    d_all = torch.from_numpy(host).to("cuda:0")
    torch.cuda.synchronize()
    lib = pwm_amd.lib()
    ctx = pwm_amd.Context(0)
    base = d_all.data_ptr()
    ptrs = (ctypes.c_void_p * T)(*[base + int(o) for o in offs])
    csz = (ctypes.c_uint64 * T)(*[int(x) for x in sizes])
    flags = pwm_amd.FLAG_TEXT_ON_DEVICE | (pwm_amd.FLAG_TREE if tree else 0)
    code = pwm_amd.ALGOS[algo]
    outs = (pwm_amd._WS * T)()

    def batched(extra):
        st = lib.pwmhip_build_batch(ctx.h, ptrs, csz, T, 1, levels, code,
                                    flags | pwm_amd.FLAG_REUSE_OUTPUT | extra, outs)
        if st != pwm_amd.OK:
            raise pwm_amd.PwmError(st, "build_batch")

    def timed(extra):
        dev = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            batched(extra)
            dev.append(ctx.last_timing())
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps * 1e3, dev

    # the launch variants are tuned on the first call; the captured graph is
    # replayed from the second on
    graph_flags = pwm_amd.FLAG_GRAPH | pwm_amd.FLAG_AUTOTUNE if graph else 0
    for _ in range(max(warmup, 1)):
        batched(graph_flags)
    batch_ms, dev = timed(graph_flags)
    dev_ms = float(np.mean([d[3] for d in dev]))
    # phase times come from the same tuned launches run eagerly (a graph is
    # timed as a whole)
    eager_ms, edev = timed(0)
    phases = [round(float(np.mean([d[i] for d in edev])), 4) for i in range(3)]
    batched(graph_flags)  # the checked structures come from a graph replay

    # one plain build per text, each into its own reused output
    singles = (pwm_amd._WS * T)()
    sflags = flags | pwm_amd.FLAG_REUSE_OUTPUT

    def per_text_builds(idx):
        devsum = 0.0
        arr = (ctypes.c_double * 4)()
        for i in idx:
            st = lib.pwmhip_build(ctx.h, ptrs[i], csz[i], 1, levels, code, sflags, ctypes.byref(singles[i]))
            if st != pwm_amd.OK:
                raise pwm_amd.PwmError(st, "build")
            lib.pwmhip_last_timing(ctx.h, arr)
            devsum += arr[3]
        return devsum

    rng = np.random.default_rng(1)
    check = sorted(set(rng.integers(0, T, 200).tolist()) | {int(np.argmax(sizes)), int(np.argmin(sizes))})
    single_ms = single_dev_ms = None
    if per_text:
        per_text_builds(range(T))  # allocates every output once
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        single_dev_ms = per_text_builds(range(T))
        torch.cuda.synchronize()
        single_ms = (time.perf_counter() - t0) * 1e3
    else:
        per_text_builds(check)

    # bit-exact: the batched structures equal the per-text ones and the
    # oracle's (wx_naive restated; zeros[levels-1] is 0 for ppc)
    if check_oracle:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
    mism = []
    for i in check:
        a, b = outs[i], singles[i]
        ok = a.levels == b.levels and list(a.zeros[:levels]) == list(b.zeros[:levels])
        t_i = host[offs[i]:offs[i] + sizes[i]]
        o = oracle_lib.naive(t_i, levels, tree) if check_oracle else None
        if o is not None and not tree:
            z = o.zeros.copy()
            if algo in ("ppc", "ppc_ss", "pc_ss", "dd_pc_ss"):
                z[-1] = 0
            ok = ok and list(a.zeros[:levels]) == [int(x) for x in z]
        for l in range(a.levels if ok else 0):
            nw = (int(a.level_bits[l]) + 63) // 64
            x = np.empty(nw, np.uint64)
            y = np.empty(nw, np.uint64)
            lib.pwmhip_ws_download(ctx.h, ctypes.byref(a), l, x.ctypes.data_as(ctypes.c_void_p))
            lib.pwmhip_ws_download(ctx.h, ctypes.byref(b), l, y.ctypes.data_as(ctypes.c_void_p))
            ok = ok and np.array_equal(x, y) and (o is None or np.array_equal(x, o.levels[l]))
        if not ok:
            mism.append(i)

    words = ((sizes + 63) // 64).sum()
    b_min = total + levels * 8 * int(words)
    figures = {
        "texts": T, "symbols": total, "levels": levels, "algorithm": ("wt_" if tree else "wm_") + algo,
        "batched_ms": round(batch_ms, 4),
        "batched_device_ms": round(dev_ms, 4),
        "graph": graph,
        "eager_ms": round(eager_ms, 4),
        "device_ms_phases": {"local_build": phases[0], "offsets": phases[1], "stitch": phases[2]},
        "batched_input_GBps": round(total / (batch_ms / 1e3) / 1e9, 2),
        "per_text_ms": round(single_ms, 2) if single_ms else None,
        "per_text_device_ms": round(single_dev_ms, 2) if single_dev_ms else None,
        "speedup_vs_per_text": round(single_ms / batch_ms, 2) if single_ms else None,
        "checked_texts": len(check),
        "verified": not mism,
        "checked_against": "per-text builds%s, every level word and zeros" % (" and the oracle" if check_oracle else ""),
        "generate_s": round(gen_s, 2),
    }
    if mism:
        figures["mismatch_texts"] = mism[:20]
    for i in range(T):
        lib.pwmhip_ws_free(ctx.h, ctypes.byref(singles[i]))
    for i in range(T - 1, -1, -1):
        lib.pwmhip_ws_free(ctx.h, ctypes.byref(outs[i]))
    ctx.close()
    del d_all
    torch.cuda.empty_cache()
    return figures, b_min, phases, dev_ms, (host, offs, sizes)


def run_batch(args):
    """The many-texts workload as its own line."""
    seed, T, median, sigma, lo, hi, levels, algo, tree = BATCH_CONFIGS[args.config]
    fig, b_min, phases, dev_ms, (host, offs, sizes) = batch_figures(
        args.config, args.steps, args.warmup, graph=not args.no_graph, per_text=not args.no_per_text)
    roof, whole = rooflines(b_min, phases, dev_ms, args.config)
    fig["build_roofline"] = whole
    fig["workload"] = ("%s: %d uint8 DNA texts, lengths lognormal(median %d, sigma %.1f) clipped [%d, %d], "
                       "%d symbols total, %d levels, %s %s, one batched build vs one build per text"
                       % (args.config, T, median, sigma, lo, hi, fig["symbols"], levels, "wt" if tree else "wm",
                          algo))
    out = {
        "metric": "WM/WT build input GB/s + % HBM peak; batched vs per-text launches, bit-exact",
        "value": fig["batched_input_GBps"],
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fig["batched_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (dna generator, seed 0x%X; lognormal lengths)" % seed,
        "verified": fig["verified"],
        "config": fig,
        "roofline": roof,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_batch(host, offs, sizes, levels, tree, args.cpu_sample_mib << 20)
    print(json.dumps(out), flush=True)
    if not fig["verified"]:
        sys.exit(1)


# Serving: Poisson arrivals of build+query requests (seed, requests per rate,
# median length, sigma, min length, max length, queries per request, budget)
SERVE_CONFIGS = {
    "serve": (0xC5, 200, 1 << 20, 2.0, 4 << 10, 256 << 20, 100_000, 200 * 10**9),
}


def run_serve(args):
    """Serving loop: saturation throughput, then latency at fractions of it."""
    import threading
    import torch
    import pwm_amd

    seed, nreq, median, sigma, lo, hi, nq, budget = SERVE_CONFIGS[args.config]
    rng = np.random.default_rng(seed)
    pool_len = 512 << 20
    # host text pools in pinned memory, one per alphabet (5, 20, 256 symbols)
    pools = []
    for kind, sig in (("dna", 5), ("u20", 20), ("u256", 256)):
        buf = torch.empty(pool_len, dtype=torch.uint8, pin_memory=True)
        if kind == "dna":
            buf.numpy()[:] = pwm_amd.generate("dna", seed, pool_len)
        else:
            buf.numpy()[:] = np.random.default_rng(seed + sig).integers(0, sig, pool_len, dtype=np.uint8)
        pools.append((buf, sig, {5: 3, 20: 5, 256: 8}[sig]))

    def make_requests(r):
        reqs = []
        for _ in range(nreq):
            buf, sig, L = pools[r.choice(3, p=[0.6, 0.3, 0.1])]
            n = int(np.clip(r.lognormal(np.log(median), sigma), lo, hi))
            off = int(r.integers(0, pool_len - n + 1))
            text = buf.numpy()[off:off + n]
            op = r.integers(0, 3, nq).astype(np.uint32)
            pos = r.integers(0, n, nq)
            a = np.where(op == 0, pos, text[r.integers(0, n, nq)]).astype(np.uint64)
            b = np.where(op == 1, r.integers(0, n + 1, nq), r.integers(1, 4, nq)).astype(np.uint64)
            reqs.append((text, L, op, a, b))
        return reqs

    def run(reqs, gaps):
        srv = pwm_amd.Server(hbm_budget_bytes=budget, workers=3)
        ids = []

        def submit():
            for (text, L, op, a, b), g in zip(reqs, gaps):
                if g > 0:
                    time.sleep(g)
                bid = srv.build(text, L, "ppc")
                ids.append((bid, srv.query(bid, op, a, b)))

        t0 = time.perf_counter()
        th = threading.Thread(target=submit)
        th.start()
        th.join()
        for bid, qid in ids:
            srv.wait(qid)
        wall = time.perf_counter() - t0
        for bid, _ in ids:
            srv.release(bid)
        st = srv.stats()
        srv.close()
        return wall, st

    reqs = make_requests(rng)
    total_bytes = sum(r[0].size for r in reqs)
    run(reqs[:20], [0.0] * 20)  # warm-up
    wall, st = run(reqs, [0.0] * nreq)  # all at once: saturation
    sat_rps = nreq / wall
    sweep = []
    for frac in (0.25, 0.5, 0.75, 1.0):
        lam = frac * sat_rps
        gaps = rng.exponential(1.0 / lam, nreq)
        w, s2 = run(reqs, gaps)
        sweep.append({"offered_rps": round(lam, 1), "achieved_rps": round(nreq / w, 1),
                      "request_p50_ms": round(s2["query_p50_ms"], 3), "request_p99_ms": round(s2["query_p99_ms"], 3),
                      "build_p50_ms": round(s2["build_p50_ms"], 3), "build_p99_ms": round(s2["build_p99_ms"], 3),
                      "texts_per_launch": round(s2["mean_texts_per_launch"], 2),
                      "peak_budget_bytes": int(s2["bytes_peak"])})
    out = {
        "metric": "WM/WT build input GB/s + % HBM peak; batched vs per-text launches, bit-exact",
        "value": round(total_bytes / wall / 1e9, 3),
        "unit": "GB/s",
        "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": round(wall * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
        "data": "synthetic (dna / uniform-20 / uniform-256 pools, seed 0x%X, pinned host memory)" % seed,
        "config": {"workload": "%s: %d requests, lengths lognormal(median %d, sigma %.1f) clipped [%d, %d], "
                               "alphabets 5/20/256 (p 0.6/0.3/0.1; levels 3/5/8), wm_ppc, %d queries each, "
                               "HBM budget %d B, 3 worker streams" % (args.config, nreq, median, sigma, lo, hi, nq,
                                                                       budget),
                   "requests": nreq, "text_bytes": int(total_bytes),
                   "saturation_rps": round(sat_rps, 1),
                   "saturation_request_p50_ms": round(st["query_p50_ms"], 3),
                   "saturation_request_p99_ms": round(st["query_p99_ms"], 3),
                   "queries_per_s": round(nreq * nq / wall, 1),
                   "texts_per_launch": round(st["mean_texts_per_launch"], 2),
                   "sweep": sweep},
        "roofline": None,
    }
    print(json.dumps(out), flush=True)


def run_copy(args):
    """A plain copy of known size: its counter readings per byte scale the
    construction kernels' readings."""
    import torch
    nbytes = args.copy_mib << 20
    src = torch.ones(nbytes // 8, dtype=torch.int64, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(args.warmup):
        dst.copy_(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    print(json.dumps({"copy_bytes": nbytes, "ms": round(dt * 1e3, 4),
                      "GBps_read_plus_write": round(2 * nbytes / dt / 1e9, 1)}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="dna4g_wm",
                    choices=sorted(CONFIGS) + sorted(STREAM_CONFIGS) + sorted(BATCH_CONFIGS) + sorted(SERVE_CONFIGS))
    ap.add_argument("--n", type=int, default=0, help="override the text length")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--queries", type=int, default=None, help="random symbol queries to time after the build")
    ap.add_argument("--no-per-text", action="store_true",
                    help="many-texts configs: skip the one-build-per-text comparison")
    ap.add_argument("--cpu-sample-mib", type=int, default=1024)
    ap.add_argument("--no-batch", action="store_true",
                    help="headline workload: skip the many-texts (batched vs per-text) figures")
    ap.add_argument("--no-graph", action="store_true", help="batches: eager launches, default variants")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only time a plain device-to-device copy of this many MiB (calibrates the "
                         "HBM byte counters of a profiler run)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; --gpus %d is not supported "
              "(the construction does not shard across GPUs here)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.copy_mib:
        run_copy(args)
        return
    if args.config in BATCH_CONFIGS:
        run_batch(args)
        return
    if args.config in SERVE_CONFIGS:
        run_serve(args)
        return

    import torch
    import pwm_amd

    streamed = args.config in STREAM_CONFIGS
    kind, seed, n, width, levels, algo, tree = CONFIGS[STREAM_CONFIGS.get(args.config, args.config)]
    if args.n:
        n = args.n
    t0 = time.perf_counter()
    text = pwm_amd.generate(kind, seed, n, width, levels)
    gen_s = time.perf_counter() - t0
    dev = torch.device("cuda:0")
    h2d_gbps = None
    if streamed:
        # pinned host text; the build copies it in slices
        pinned = torch.empty(n * width, dtype=torch.uint8, pin_memory=True)
        pinned.numpy()[:] = text.view(np.uint8)
        src = pinned.numpy().view(text.dtype)
        d_tmp = torch.empty(n * width, dtype=torch.uint8, device=dev)
        d_tmp.copy_(pinned, non_blocking=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        d_tmp.copy_(pinned, non_blocking=True)
        torch.cuda.synchronize()
        h2d_gbps = round(n * width / (time.perf_counter() - t0) / 1e9, 2)
        del d_tmp
    else:
        src = torch.from_numpy(text).to(dev)
    torch.cuda.synchronize()
    ctx = pwm_amd.Context(0)

    ws = None
    for _ in range(args.warmup):
        ws = pwm_amd.build(src, levels, algo, tree, ctx=ctx, reuse=ws, stream=streamed)
    torch.cuda.synchronize()
    dev_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ws = pwm_amd.build(src, levels, algo, tree, ctx=ctx, reuse=ws, stream=streamed)
        dev_ms.append(ctx.last_timing())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3
    dev_total = float(np.mean([d[3] for d in dev_ms]))
    phases = [round(float(np.mean([d[i] for d in dev_ms])), 4) for i in range(3)]

    # the last build against the reference's own output on the same text
    # (outside the timed region)
    ver = verify_against_reference(args.config, text, ws, pwm_amd.lib(), ctx)

    # text read once + every level's words written once (Huffman levels are ragged)
    b_min = n * width + 8 * sum((ws.level_bit_size(l) + 63) // 64 for l in range(ws.levels()))
    passes = 1 if (levels <= 4 or algo.startswith("huff")) else (levels + 3) // 4
    roof, whole = rooflines(b_min, phases, dev_total, args.config, passes)
    value = n * width / (ms_per_step / 1e3) / 1e9
    out = {
        "metric": "WM/WT build input GB/s + % HBM peak; batched vs per-text launches, bit-exact",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u%d" % (8 * width),
        "data": "synthetic (%s generator, seed 0x%X)" % (kind, seed),
        "verified": ver["verified"],
        "config": {"workload": "%s: %d x uint%d, %d levels, %s %s" % (
            args.config, n, 8 * width, ws.levels(), "wt" if tree else "wm", algo),
            "symbols": n, "levels": ws.levels(), "algorithm": ("wt_" if tree else "wm_") + algo,
            "device_ms_per_build": round(dev_total, 4),
            "device_ms_phases": {"local_build": phases[0], "offsets": phases[1], "stitch": phases[2]},
            "symbols_per_s": round(n / (ms_per_step / 1e3), 1),
            "text": "pinned host memory, streamed in 256 Mi-symbol slices" if streamed else "resident in HBM",
            "plain_h2d_copy_GBps": h2d_gbps,
            "build_roofline": whole,
            "verification": ver,
            "generate_s": round(gen_s, 2)},
        "roofline": roof,
    }
    nq = args.queries if args.queries is not None else QUERIES.get(args.config, 0)
    if nq:
        out["config"]["queries"] = time_queries(pwm_amd, ctx, ws, text, nq)
    ws.free()
    ctx.close()
    del src
    torch.cuda.empty_cache()
    # the headline workload's line also carries the many-texts comparison
    # (batched launches against one plain launch per text, same exactness)
    if args.config in BATCH_WITH and not args.no_batch:
        bname = BATCH_WITH[args.config]
        fig = batch_figures(bname, max(3, args.steps // 2), max(1, args.warmup // 2))[0]
        fig["workload"] = bname
        out["batched"] = fig
        out["verified"] = bool(out["verified"]) and fig["verified"]
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(text, levels, tree, width, args.cpu_sample_mib << 20, algo)
    print(json.dumps(out), flush=True)
    if out["verified"] is False:
        sys.exit(1)


def time_queries(pwm_amd, ctx, ws, text, nq, checked=20000):
    """Random access / rank / select over the built structure: throughput of
    one batch of nq queries with device-resident arguments, and a bit-exact
    check of a sample against positions computed on the host."""
    import ctypes
    import torch
    rng = np.random.default_rng(0xC3)
    n = text.size
    qs = pwm_amd.QuerySupport(ws)
    op = rng.integers(0, 3, nq).astype(np.uint32)
    pos = rng.integers(0, n, nq)
    sym = text[rng.integers(0, n, nq)].astype(np.uint64)
    a = np.where(op == 0, pos, sym).astype(np.uint64)
    b = np.where(op == 1, rng.integers(0, n + 1, nq), rng.integers(1, 3, nq)).astype(np.uint64)
    dev = torch.device("cuda:0")
    d_op = torch.from_numpy(op.view(np.int32)).to(dev)
    d_a = torch.from_numpy(a.view(np.int64)).to(dev)
    d_b = torch.from_numpy(b.view(np.int64)).to(dev)
    d_r = torch.empty(nq, dtype=torch.int64, device=dev)
    lib = pwm_amd.lib()

    def go():
        st = lib.pwmhip_query(ctx.h, ctypes.byref(ws._ws), qs.h, ctypes.c_void_p(d_op.data_ptr()),
                              ctypes.c_void_p(d_a.data_ptr()), ctypes.c_void_p(d_b.data_ptr()), nq,
                              ctypes.c_void_p(d_r.data_ptr()), pwm_amd.FLAG_TEXT_ON_DEVICE)
        if st:
            raise pwm_amd.PwmError(st, "query")

    go()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        go()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    res = d_r.cpu().numpy().view(np.uint64)
    acc = op == 0
    ok = bool(np.array_equal(res[acc], text[a[acc]].astype(np.uint64)))
    # rank / select checked on a second batch over 16 symbols of the text,
    # against their positions found on the host
    syms = np.unique(text[rng.integers(0, n, 16)])
    occ = {int(c): np.flatnonzero(text == c) for c in syms}
    m = checked
    op2 = rng.integers(1, 3, m).astype(np.uint32)
    a2 = syms[rng.integers(0, syms.size, m)].astype(np.uint64)
    b2 = np.empty(m, np.uint64)
    want = np.empty(m, np.uint64)
    for k in range(m):
        p_ = occ[int(a2[k])]
        if op2[k] == 1:
            b2[k] = rng.integers(0, n + 1)
            want[k] = np.searchsorted(p_, b2[k], "left")
        else:
            b2[k] = rng.integers(1, p_.size + 2)
            want[k] = p_[b2[k] - 1] if b2[k] <= p_.size else np.uint64(2**64 - 1)
    got = qs.query(op2, a2, b2)
    ok = ok and bool(np.array_equal(got, want))
    qs.free()
    return {"count": nq, "ms": round(dt * 1e3, 3), "queries_per_s": round(nq / dt, 1),
            "mix": "access/rank/select uniform; positions uniform; symbols drawn from the text",
            "checked": "%d access in the timed batch + %d rank/select over 16 symbols" % (int(acc.sum()), m),
            "bit_exact": ok}


if __name__ == "__main__":
    main()
