"""Job traces for the sort / select / histogram server: a line-oriented JSON
format for jobs, results and run statistics, so that a load can be written
once, replayed against the server (pa_server_*) and two runs compared.

The reference has no server; its nearest ideas are bundling small work items
(BundleSmallWorkItems.cpp:36-144) and choosing an algorithm by free memory
(MemoryUsage.cpp:20-86).  The job fields follow the C ABI's pa_job
(include/pa_hip.h).

Job line (one JSON object per line):
    {"id": 7, "t_us": 1250.0, "op": "select", "dtype": "u64", "n": 1048576,
     "k": 1000, "shift": 0, "bits": 8, "seed": 1240007}
  op     "sort" (in place, ascending), "select" (k-th smallest, 0-based) or
         "hist" (2^bits counts of (key >> shift) & (2^bits - 1))
  dtype  "u32" or "u64"
  t_us   arrival time after the start of the run, microseconds
  seed   input keys are pa_fill_random_{u32,u64}(seed): the same seed gives
         the same keys on every run and every device
Result line:
    {"id": 7, "status": "ok", "latency_ms": 0.84, "digest": "9f2c...", "n": ...}
  digest fixes the job's output bit for bit: for a sort, the sum of the keys
  and a position-weighted sum (both mod 2^64), for select the value, for a
  histogram the counts; equal digests on two runs mean equal outputs.
Stats line (one JSON object):
    {"jobs", "span_s", "jobs_per_s", "keys", "gkeys_per_s", "p50_ms",
     "p90_ms", "p99_ms", "max_ms", "launches", "coalesced_jobs",
     "peak_bytes_in_flight"}

Command line:
    python -m parallelalgorithms_amd.trace generate OUT --jobs N [--seed S] [--rate JOBS_PER_S]
    python -m parallelalgorithms_amd.trace replay TRACE [--results OUT] [--streams K] [--budget-gb G]
    python -m parallelalgorithms_amd.trace compare RESULTS_A RESULTS_B
"""
import argparse
import json
import sys
from dataclasses import asdict, dataclass

OPS = ("sort", "select", "hist")
DTYPES = ("u32", "u64")
_MASK64 = (1 << 64) - 1


@dataclass
class Job:
    id: int
    t_us: float
    op: str
    dtype: str
    n: int
    k: int = 0
    shift: int = 0
    bits: int = 8
    seed: int = 0

    def validate(self):
        if self.op not in OPS:
            raise ValueError(f"job {self.id}: op must be one of {OPS}, not {self.op!r}")
        if self.dtype not in DTYPES:
            raise ValueError(f"job {self.id}: dtype must be one of {DTYPES}, not {self.dtype!r}")
        if self.n < 1:
            raise ValueError(f"job {self.id}: n must be >= 1")
        if self.t_us < 0:
            raise ValueError(f"job {self.id}: t_us must be >= 0")
        if self.op == "select" and not 0 <= self.k < self.n:
            raise ValueError(f"job {self.id}: select needs 0 <= k < n")
        width = 32 if self.dtype == "u32" else 64
        # pa_hist_u32 / pa_hist_u64 take 1..12 bits (at most 4096 bins).
        if self.op == "hist" and not (1 <= self.bits <= 12 and 0 <= self.shift and self.shift + self.bits <= width):
            raise ValueError(f"job {self.id}: hist needs 1 <= bits <= 12 and shift + bits <= {width}")
        return self

    def to_json(self):
        return json.dumps(asdict(self), separators=(",", ":"))

    @staticmethod
    def from_dict(d):
        known = {f for f in Job.__dataclass_fields__}
        extra = set(d) - known
        if extra:
            raise ValueError(f"unknown job fields {sorted(extra)}")
        return Job(**d).validate()


def write_jobs(path, jobs):
    with open(path, "w") as f:
        for j in jobs:
            f.write(j.validate().to_json() + "\n")


def read_jobs(path):
    jobs = []
    with open(path) as f:
        for line in f:
            if line.strip():
                jobs.append(Job.from_dict(json.loads(line)))
    ids = [j.id for j in jobs]
    if len(set(ids)) != len(ids):
        raise ValueError("job ids must be unique")
    return jobs


def write_results(path, results):
    with open(path, "w") as f:
        for r in results:
            f.write(json.dumps(r, separators=(",", ":")) + "\n")


def read_results(path):
    with open(path) as f:
        return [json.loads(line) for line in f if line.strip()]


def _unit_floats(u64):
    import numpy as np
    return (u64 >> np.uint64(11)).astype(np.float64) / float(1 << 53)


def generate(jobs=500, seed=1240, rate=None, lo=4096, hi=64 << 20, mix=(70, 20, 10)):
    """A synthetic load: sizes log-uniform in [lo, hi], op mix sort/select/hist
    in percent, u32/u64 keys 50/50, select ranks log-uniform in 1..n/10,
    top-byte histograms, Poisson arrivals at `rate` jobs/s (all at 0 if None)."""
    import numpy as np
    from . import gen_u64
    r = gen_u64(seed, 5 * jobs)
    sizes = np.exp(np.log(lo) + _unit_floats(r[:jobs]) * (np.log(hi) - np.log(lo))).astype(np.int64)
    kind = r[jobs:2 * jobs] % np.uint64(100)
    wide = (r[2 * jobs:3 * jobs] & np.uint64(1)).astype(bool)
    ku = _unit_floats(r[3 * jobs:4 * jobs])
    if rate:
        gaps = -np.log1p(-_unit_floats(r[4 * jobs:])) / rate
        t = np.cumsum(gaps) - gaps[0]
    else:
        t = np.zeros(jobs)
    out = []
    for i in range(jobs):
        n = int(sizes[i])
        op = "sort" if kind[i] < mix[0] else ("select" if kind[i] < mix[0] + mix[1] else "hist")
        dt = "u64" if wide[i] else "u32"
        k = min(max(int(np.floor(np.exp(ku[i] * np.log(max(2.0, n / 10.0))))) - 1, 0), n - 1)
        out.append(Job(id=i, t_us=round(float(t[i]) * 1e6, 3), op=op, dtype=dt, n=n, k=k if op == "select" else 0,
                       shift=(56 if dt == "u64" else 24) if op == "hist" else 0, bits=8, seed=seed * 1000 + i))
    return out


def digest_keys(t):
    """Order-sensitive digest of a device key array (int32 views of u32 keys,
    int64 views of u64 keys): sum of the keys and sum of key * (2i + 1), both
    mod 2^64, as 32 hex digits."""
    import torch
    x = t.to(torch.int64)
    if t.dtype == torch.int32:
        x = x & 0xFFFFFFFF
    w = torch.arange(x.numel(), dtype=torch.int64, device=x.device) * 2 + 1
    s1 = int(x.sum().item()) & _MASK64  # int64 sums wrap mod 2^64
    s2 = int((x * w).sum().item()) & _MASK64
    return f"{s1:016x}{s2:016x}"


def latency_stats(latencies_ms):
    import numpy as np
    lat = np.asarray(latencies_ms, dtype=np.float64)
    if lat.size == 0:
        return {"p50_ms": None, "p90_ms": None, "p99_ms": None, "max_ms": None}
    return {"p50_ms": round(float(np.percentile(lat, 50)), 3), "p90_ms": round(float(np.percentile(lat, 90)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3), "max_ms": round(float(lat.max()), 3)}


def _prepare(jobs):
    """Device inputs (generated from each job's seed) and output buffers."""
    import torch
    from . import fill_random_u32, fill_random_u64
    arrs, outs = [], []
    for j in jobs:
        j.validate()
        wide = j.dtype == "u64"
        t = torch.empty(j.n, dtype=torch.int64 if wide else torch.int32, device="cuda")
        (fill_random_u64 if wide else fill_random_u32)(t, j.seed)
        arrs.append(t)
        if j.op == "select":
            outs.append(torch.zeros(1, dtype=torch.int64 if wide else torch.int32, device="cuda"))
        elif j.op == "hist":
            outs.append(torch.zeros(1 << j.bits, dtype=torch.int64, device="cuda"))
        else:
            outs.append(None)
    torch.cuda.synchronize()
    return arrs, outs


def _add_digests(jobs, results, arrs, outs):
    for i, j in enumerate(jobs):
        r = results[i]
        if r["status"] != "ok":
            continue
        if j.op == "sort":
            r["digest"] = digest_keys(arrs[i])
        elif j.op == "select":
            v = int(outs[i].cpu()[0]) & (_MASK64 if j.dtype == "u64" else 0xFFFFFFFF)
            r["digest"] = f"{v:x}"
        else:
            r["digest"] = ",".join(str(int(c)) for c in outs[i].cpu().tolist())


def replay_plain(jobs, digests=True):
    """The same jobs as plain library calls, one after another in arrival
    order on the default stream, with no server: the per-job baseline the
    server's throughput is compared with.  Returns (results, stats) like
    replay; latencies are not defined here."""
    import time
    import torch
    from . import hist_u32, hist_u64, select_u32, select_u64, sort_u32, sort_u64
    arrs, outs = _prepare(jobs)
    results = []
    t0 = time.perf_counter()
    for i, j in enumerate(jobs):
        wide = j.dtype == "u64"
        try:
            if j.op == "sort":
                (sort_u64 if wide else sort_u32)(arrs[i])
            elif j.op == "select":
                (select_u64 if wide else select_u32)(arrs[i], j.k, outs[i])
            else:
                (hist_u64 if wide else hist_u32)(arrs[i], j.shift, j.bits, outs[i])
            results.append({"id": j.id, "status": "ok", "latency_ms": None})
        except Exception as e:
            results.append({"id": j.id, "status": f"error: {e}", "latency_ms": None})
    torch.cuda.synchronize()
    span = time.perf_counter() - t0
    if digests:
        _add_digests(jobs, results, arrs, outs)
    keys = float(sum(j.n for j in jobs))
    stats = {"jobs": len(jobs), "span_s": round(span, 6), "jobs_per_s": round(len(jobs) / span, 1),
             "keys": int(keys), "gkeys_per_s": round(keys / span / 1e9, 3),
             "failed": sum(r["status"] != "ok" for r in results)}
    del arrs, outs
    torch.cuda.empty_cache()
    return results, stats


def replay(jobs, budget_bytes=64 << 30, streams=4, coalesce_max_keys=1 << 20, paced=True, digests=True):
    """Runs the jobs through one server, each submitted at its t_us (or all at
    once when paced is False).  Inputs are generated on the device before the
    clock starts.  Returns (result dicts in job order, stats dict)."""
    import time
    import torch
    from . import KEY_U32, KEY_U64, OP_HIST, OP_SELECT, OP_SORT, Server
    opc = {"sort": OP_SORT, "select": OP_SELECT, "hist": OP_HIST}
    arrs, outs = _prepare(jobs)
    srv = Server(budget_bytes=budget_bytes, streams=streams, coalesce_max_keys=coalesce_max_keys)
    results = []
    try:
        tickets = []
        t0 = time.perf_counter()
        for i, j in enumerate(jobs):
            if paced:
                due = t0 + j.t_us / 1e6
                if due - time.perf_counter() > 2e-3:
                    time.sleep(due - time.perf_counter() - 1e-3)
                while time.perf_counter() < due:
                    pass
            tickets.append(srv.submit(opc[j.op], arrs[i], KEY_U64 if j.dtype == "u64" else KEY_U32, k=j.k,
                                      shift=j.shift, bits=j.bits, out=outs[i]))
        lats = []
        for i, j in enumerate(jobs):
            try:
                lat = srv.wait(tickets[i])
                results.append({"id": j.id, "status": "ok", "latency_ms": round(lat, 4)})
                lats.append(lat)
            except Exception as e:  # the job's own failure; the run goes on
                results.append({"id": j.id, "status": f"error: {e}", "latency_ms": None})
        span = time.perf_counter() - t0
        st = srv.stats()
    finally:
        srv.close()
    torch.cuda.synchronize()
    if digests:
        _add_digests(jobs, results, arrs, outs)
    keys = float(sum(j.n for j in jobs))
    stats = {"jobs": len(jobs), "span_s": round(span, 6), "jobs_per_s": round(len(jobs) / span, 1),
             "keys": int(keys), "gkeys_per_s": round(keys / span / 1e9, 3), **latency_stats(lats),
             "failed": sum(r["status"] != "ok" for r in results), "launches": st["launches"],
             "coalesced_jobs": st["coalesced_jobs"], "peak_bytes_in_flight": st["peak_bytes_in_flight"],
             "admission_waits": st["admission_waits"], "scratch_releases": st["scratch_releases"],
             "oversize_units": st["oversize_units"]}
    del arrs, outs
    torch.cuda.empty_cache()
    return results, stats


def compare(results_a, results_b):
    """Jobs whose outputs differ between two runs (by digest), jobs missing
    from either, and the latency percentiles of both."""
    a = {r["id"]: r for r in results_a}
    b = {r["id"]: r for r in results_b}
    common = sorted(set(a) & set(b))
    differ = [i for i in common if a[i].get("digest") != b[i].get("digest") or a[i]["status"] != b[i]["status"]]
    return {"jobs": len(common), "outputs_differ": differ, "only_in_a": sorted(set(a) - set(b)),
            "only_in_b": sorted(set(b) - set(a)),
            "a": latency_stats([a[i]["latency_ms"] for i in common if a[i].get("latency_ms") is not None]),
            "b": latency_stats([b[i]["latency_ms"] for i in common if b[i].get("latency_ms") is not None])}


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m parallelalgorithms_amd.trace")
    sub = ap.add_subparsers(dest="cmd", required=True)
    g = sub.add_parser("generate")
    g.add_argument("out")
    g.add_argument("--jobs", type=int, default=500)
    g.add_argument("--seed", type=int, default=1240)
    g.add_argument("--rate", type=float, default=None, help="Poisson arrivals, jobs/s (default: all at t=0)")
    g.add_argument("--min-n", type=int, default=4096)
    g.add_argument("--max-n", type=int, default=64 << 20)
    r = sub.add_parser("replay")
    r.add_argument("trace")
    r.add_argument("--results")
    r.add_argument("--streams", type=int, default=4)
    r.add_argument("--budget-gb", type=float, default=64)
    r.add_argument("--coalesce-max-keys", type=int, default=1 << 20)
    r.add_argument("--unpaced", action="store_true", help="submit every job at once")
    c = sub.add_parser("compare")
    c.add_argument("a")
    c.add_argument("b")
    args = ap.parse_args(argv)
    if args.cmd == "generate":
        write_jobs(args.out, generate(args.jobs, args.seed, args.rate, args.min_n, args.max_n))
        return 0
    if args.cmd == "replay":
        results, stats = replay(read_jobs(args.trace), int(args.budget_gb * (1 << 30)), args.streams,
                                args.coalesce_max_keys, paced=not args.unpaced)
        if args.results:
            write_results(args.results, results)
        print(json.dumps(stats), flush=True)
        return 1 if stats["failed"] else 0
    rep = compare(read_results(args.a), read_results(args.b))
    print(json.dumps(rep), flush=True)
    return 1 if rep["outputs_differ"] or rep["only_in_a"] or rep["only_in_b"] else 0


if __name__ == "__main__":
    sys.exit(main())
