#!/usr/bin/env python3
"""Benchmark: LSD radix sort of 100,000,000 uniform u32 keys on one MI355X.

One step = one full sort (fused digit histogram + 4 onesweep passes of 8-bit
digits) of a fresh device-resident copy of the input, keys drawn from
std::mt19937_64(1234) truncated to 32 bits (the reference's generator,
ParallelAlgorithms.cpp:91,114-115).  Prints ONE JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--no-cpu]

roofline: the dominant kernel is k_onesweep; its algorithmic bytes per launch
are n x (4 B read + 4 B write) and its average duration is measured with HIP
events recorded on the launch stream around every launch in the timed region.
cpu_baseline: the oracle's restatement of the reference's SortRadixPar
(RadixSortLsdParallel.h:367-387) on this host's cores, same input.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
REFERENCE_MKEYS = 2865.0  # README.md:9, SortRadixPar, 100M u32 random, 48-core AWS C7a


def cpu_threads():
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


def cpu_baseline(keys_np, iters=3):
    """Times the oracle's SortRadixPar restatement (3-arg form) on host cores."""
    import numpy as np
    import oracle
    threads = cpu_threads()
    oracle.set_threads(threads)
    lib = oracle.load()
    a = np.empty_like(keys_np)
    tmp = np.empty_like(keys_np)
    times = []
    for it in range(iters + 1):
        a[:] = keys_np
        tmp[:] = 0  # page in the scratch buffer (RadixSortLsdBenchmark.cpp:86-91)
        t0 = time.perf_counter()
        lib.orc_sort_radix_par_u32(a.ctypes.data, tmp.ctypes.data, a.size, 512 * 1024)
        t1 = time.perf_counter()
        if it > 0:
            times.append(t1 - t0)
    ok = bool(np.all(a[1:] >= a[:-1]))
    med = statistics.median(times)
    return {
        "value": round(keys_np.size / med / 1e9, 4),
        "unit": "Gkeys/s",
        "cores": threads,
        "kind": "port",
        "sample": f"full workload: {keys_np.size} u32 keys, median of {iters} runs of the oracle's SortRadixPar "
                  f"restatement (std::thread pool), sorted={ok}",
    }


def load_traffic(kernel, workload):
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get("traffic_bytes_per_launch")
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=100_000_000)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true", help="skip per-kernel event timing")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is described in DESIGN.md, not built)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import parallelalgorithms_amd as pa

    pa.load()
    dev = torch.device("cuda:0")
    props = torch.cuda.get_device_properties(dev)
    n = args.n
    print(f"[bench] device={props.name} CUs={props.multi_processor_count} n={n}", file=sys.stderr, flush=True)

    keys_np = pa.gen_u32(args.seed, n)
    src = torch.from_numpy(keys_np.view(np.int32)).to(dev)
    tmp = torch.empty_like(src)
    stream = torch.cuda.current_stream()

    # Achievable-bandwidth probe: plain copy kernel of the same array.
    cp = torch.empty_like(src)
    for _ in range(3):
        pa.copy_u32(cp, src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(10):
        pa.copy_u32(cp, src)
    e1.record(stream)
    torch.cuda.synchronize()
    copy_gbs = 10 * 2 * 4 * n / (e0.elapsed_time(e1) / 1e3) / 1e9
    del cp

    # Warmup on scratch copies.
    for _ in range(args.warmup):
        w = src.clone()
        pa.sort_u32(w, tmp)
    torch.cuda.synchronize()
    ref = torch.sort(src.to(torch.int64) & 0xFFFFFFFF)[0].to(torch.int32) if n <= 200_000_000 else None
    if ref is not None:
        ok = bool(torch.equal(w, ref))
        del ref
    else:
        ok = None
    print(f"[bench] warmup done, sorted output equals torch.sort: {ok}", file=sys.stderr, flush=True)

    bufs = [src.clone() for _ in range(args.steps)]
    torch.cuda.synchronize()
    if not args.no_profile:
        pa.profile_reset()
        pa.profile_enable(True)
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    for b in bufs:
        pa.sort_u32(b, tmp)
    stop.record(stream)
    torch.cuda.synchronize()
    pa.profile_enable(False)
    ms_total = start.elapsed_time(stop)
    ms_step = ms_total / args.steps
    value = n / (ms_step / 1e3) / 1e9

    roofline = None
    kernels = {}
    if not args.no_profile:
        for k in ("k_sort_prepare", "k_hist_digits", "k_onesweep"):
            tot, cnt = pa.profile_query(k)
            if cnt:
                kernels[k] = {"avg_us": round(tot / cnt * 1e3, 2), "launches": cnt}
        if "k_onesweep" in kernels:
            avg_s = kernels["k_onesweep"]["avg_us"] / 1e6
            alg = 8.0 * n  # 4 B read + 4 B write per key per digit pass
            achieved = alg / avg_s / 1e9
            roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "traffic": load_traffic("k_onesweep", f"{n // 1_000_000}M_u32"),
                        "kernel": "k_onesweep", "algorithmic_bytes_per_launch": int(alg)}
        pa.profile_reset()
    del bufs

    cpu = None
    if not args.no_cpu:
        print("[bench] timing CPU baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(keys_np)

    whole_sort_bytes = 36.0 * n  # 1 histogram read + 4 x (read + write)
    out = {
        "metric": "Gkeys/s LSD sort of 100M u32 (uniform, mt19937_64 seed 1234)",
        "value": round(value, 3),
        "unit": "Gkeys/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(value / (REFERENCE_MKEYS / 1e3), 2),
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"lsd_sort_u32_uniform_{n}", "n": n, "key_bits": 32, "digit_bits": 8, "passes": 4,
                   "keys": "mt19937_64(1234) truncated to u32", "device_resident": True},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "whole_sort_hbm_frac": round(whole_sort_bytes / (ms_step / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        "copy_kernel_GBps": round(copy_gbs, 1),
        "kernels": kernels,
        "sorted_ok": ok,
        "baseline_ref": "README.md:9 SortRadixPar 100M u32 random, 48-core AWS C7a: 2865 Mkeys/s",
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
