"""Benchmark: wavelet matrix / tree construction throughput on one MI355X.

Default workload (the headline configuration): a 4 GiB uint8 DNA-like text
(codes 0..4, order-2 Markov source, N runs), 3 levels, wavelet matrix built
with the prefix-counting algorithm's drop-in (wm_ppc).  A step is one full
construction of all levels from the text resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config NAME]

Prints one JSON line.  `value` is input GB/s (n * bytes per symbol / time per
step); `roofline` relates the minimal algorithmic HBM bytes of a build
(text read once + every level written once) to the device time of the
build's kernels; `cpu_baseline` is the reference's own OpenMP algorithms
(compiled from its headers, oracle/_ref) on a bounded prefix of the same
text, on this host's cores.
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
}


def cpu_baseline(text, levels, tree, width, sample_bytes):
    """Reference algorithms (wx_ppc / wx_pps / wx_dd_pc) on a prefix sample."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    cores = len(os.sched_getaffinity(0))
    sample = np.ascontiguousarray(text[: sample_bytes // width])
    if oracle_lib.have_ref():
        os.environ.setdefault("OMP_NUM_THREADS", str(cores))
        best = None
        per = {}
        for algo in ("ppc", "pps", "dd_pc"):
            oracle_lib.ref_build(algo, sample, levels, tree)  # warm-up
            ts = sorted(oracle_lib.ref_build(algo, sample, levels, tree)[1] for _ in range(3))
            per[algo] = round(sample.nbytes / ts[1] / 1e9, 3)
            if best is None or per[algo] > best[1]:
                best = (algo, per[algo])
        return {"value": best[1], "unit": "GB/s", "cores": int(os.environ["OMP_NUM_THREADS"]),
                "kind": "reference",
                "sample": "%d MiB prefix of the same text; wx_%s best of ppc/pps/dd_pc (median of 3): %s"
                          % (sample.nbytes >> 20, best[0], per)}
    small = sample[: 16 << 20]
    t0 = time.perf_counter()
    oracle_lib.naive(small, levels, tree)
    dt = time.perf_counter() - t0
    return {"value": round(small.nbytes / dt / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": "%d MiB prefix, scalar CPU restatement (reference build absent)" % (small.nbytes >> 20)}


def traffic_from_profiles(config):
    p = os.path.join(ROOT, "profiles", "%s_traffic.json" % config)
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f).get("bytes_per_build")
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="dna4g_wm", choices=sorted(CONFIGS))
    ap.add_argument("--n", type=int, default=0, help="override the text length")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample-mib", type=int, default=256)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; --gpus %d is not supported "
              "(the construction does not shard across GPUs here)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    import pwm_amd

    kind, seed, n, width, levels, algo, tree = CONFIGS[args.config]
    if args.n:
        n = args.n
    t0 = time.perf_counter()
    text = pwm_amd.generate(kind, seed, n, width, levels)
    gen_s = time.perf_counter() - t0
    dev = torch.device("cuda:0")
    d_text = torch.from_numpy(text).to(dev)
    torch.cuda.synchronize()
    ctx = pwm_amd.Context(0)

    ws = None
    for _ in range(args.warmup):
        ws = pwm_amd.build(d_text, levels, algo, tree, ctx=ctx, reuse=ws)
    torch.cuda.synchronize()
    dev_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ws = pwm_amd.build(d_text, levels, algo, tree, ctx=ctx, reuse=ws)
        dev_ms.append(ctx.last_timing())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3
    dev_total = float(np.mean([d[3] for d in dev_ms]))
    phases = [round(float(np.mean([d[i] for d in dev_ms])), 4) for i in range(3)]

    # light self-check of the last build: zeros against a host histogram
    z = ws.zeros()
    if not tree:
        hist = np.bincount(text[: min(n, 1 << 28)], minlength=1 << levels) if width == 1 and n <= (1 << 28) else None
        if hist is not None:
            exp0 = int(hist[: 1 << (levels - 1)].sum())
            assert int(z[0]) == exp0, "zeros[0] mismatch"

    b_min = n * width + levels * 8 * ((n + 63) // 64)
    achieved = b_min / (dev_total / 1e3) / 1e9
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
        "config": {"workload": "%s: %d x uint%d, %d levels, %s %s" % (
            args.config, n, 8 * width, levels, "wt" if tree else "wm", algo),
            "symbols": n, "levels": levels, "algorithm": ("wt_" if tree else "wm_") + algo,
            "device_ms_per_build": round(dev_total, 4),
            "device_ms_phases": {"local_build": phases[0], "offsets": phases[1], "stitch": phases[2]},
            "symbols_per_s": round(n / (ms_per_step / 1e3), 1),
            "generate_s": round(gen_s, 2)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profiles(args.config)},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(text, levels, tree, width, args.cpu_sample_mib << 20)
    ws.free()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
