"""The server's admission against the library's own scratch.

A job mix of about 2 GB of scratch served under a 256 MB budget must keep the
scratch the server's streams hold within the budget, must take the wait path,
and must give every job the result of the same call made alone; a job that
fails (a 13-bit histogram) returns its status while the others complete.  A
job larger than the whole budget runs alone.  The charge per job
(pa_job_scratch_bytes) is checked against what the library then holds.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

M = 1 << 20


def _keys(P, kind, n, seed):
    import torch
    if kind.endswith("64"):
        t = torch.empty(n, dtype=torch.int64, device="cuda")
        P.fill_random_u64(t, seed)
    else:
        t = torch.empty(n, dtype=torch.int32, device="cuda")
        P.fill_random_u32(t, seed)
    return t


def _mix(P):
    """(kind, n, k, shift, bits) jobs; sizes cycle so streams' blocks grow.
    Four 20M-key u64 sorts (about 200 MiB of scratch each, under 1 ms of GPU
    time) come in pairs: the second of a pair cannot fit beside the first's
    block while the first runs, so the dispatcher waits."""
    jobs = [("sort64", 20 * M, 0, 0, 0), ("sort64", 20 * M, 0, 0, 0)]
    for i in range(8):
        jobs.append(("sort32", (8, 16, 24)[i % 3] * M, 0, 0, 0))
    for i in range(4):
        jobs.append(("sort64", (4, 8, 12)[i % 3] * M, 0, 0, 0))
    jobs += [("sort64", 20 * M, 0, 0, 0), ("sort64", 20 * M, 0, 0, 0)]
    for i in range(8):
        n = (16, 32)[i % 2] * M
        jobs.append(("select32", n, (n // 2, n // 100, 0, n - 1)[i % 4], 0, 0))
    for i in range(4):
        n = (8, 16)[i % 2] * M
        jobs.append(("select64", n, (n // 3, n // 100)[i % 2], 0, 0))
    for i in range(6):
        jobs.append(("hist64" if i % 2 else "hist32", 16 * M, 0, (20, 40)[i % 2] if i % 2 else 20, 4 + 2 * (i % 5)))
    r = P.gen_u64(2026, 40)
    for i in range(40):
        jobs.append(("sort32", 100 + int(r[i] % 60000), 0, 0, 0))  # small: coalesced
    # Interleave big and small so batches form between large jobs.
    big, small = jobs[:36], jobs[36:]
    out = []
    for i in range(max(len(big), len(small))):
        if i < len(big):
            out.append(big[i])
        if i < len(small):
            out.append(small[i])
    return out


def _submit(P, srv, kind, d, k, shift, bits):
    import torch
    out = None
    if kind.startswith("select"):
        out = torch.empty(1, dtype=torch.int64 if kind.endswith("64") else torch.int32, device="cuda")
        t = srv.submit(P.OP_SELECT, d, P.KEY_U64 if kind.endswith("64") else P.KEY_U32, k=k, out=out)
    elif kind.startswith("hist"):
        out = torch.empty(1 << min(bits, 12), dtype=torch.int64, device="cuda")
        t = srv.submit(P.OP_HIST, d, P.KEY_U64 if kind.endswith("64") else P.KEY_U32, shift=shift, bits=bits, out=out)
    else:
        t = srv.submit(P.OP_SORT, d, P.KEY_U64 if kind == "sort64" else P.KEY_U32)
    return t, out


def _solo(P, kind, src, k, shift, bits):
    """The same call made alone (default stream, no server) on a copy."""
    import torch
    x = src.clone()
    if kind == "sort32":
        P.sort_u32(x)
        return x
    if kind == "sort64":
        P.sort_u64(x)
        return x
    if kind.startswith("select"):
        r = torch.empty(1, dtype=x.dtype, device="cuda")
        (P.select_u64 if kind.endswith("64") else P.select_u32)(x, k, r)
        return r
    c = torch.empty(1 << bits, dtype=torch.int64, device="cuda")
    (P.hist_u64 if kind.endswith("64") else P.hist_u32)(x, shift, bits, c)
    return c


def _as_i64_order(t, kind):
    """Keys as int64 values whose signed order is the unsigned key order."""
    import torch
    if kind.endswith("64"):
        return t ^ torch.tensor(-(1 << 63), dtype=torch.int64, device=t.device)
    return t.to(torch.int64) & 0xFFFFFFFF


def _charge(P, kind, n, bits):
    op = P.OP_SORT if kind.startswith("sort") else P.OP_SELECT if kind.startswith("select") else P.OP_HIST
    return P.Server.job_scratch_bytes(op, n, P.KEY_U64 if kind.endswith("64") else P.KEY_U32, bits)


def test_admission_budget_binds(pa):
    import torch
    P = pa
    budget = 256 * M
    mix = _mix(P)
    charged = sum(_charge(P, kind, n, bits) for kind, n, _, _, bits in mix)
    assert charged > 1800 * M  # the mix asks for about 2 GB of scratch
    assert max(_charge(P, kind, n, bits) for kind, n, _, _, bits in mix) <= budget  # no oversize unit here
    srcs = [_keys(P, kind, n, 700 + i) for i, (kind, n, _, _, _) in enumerate(mix)]
    dev = [s.clone() for s in srcs]
    bad_keys = _keys(P, "hist32", 1 * M, 99)
    bad_out = torch.empty(1 << 12, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    print(f"admission: {len(mix)} jobs, {charged / M:.0f} MiB charged, budget {budget / M:.0f} MiB", flush=True)
    srv = P.Server(budget_bytes=budget, streams=4, coalesce_max_keys=1 << 16)
    try:
        tickets, outs = [], []
        for i, ((kind, n, k, shift, bits), d) in enumerate(zip(mix, dev)):
            t, out = _submit(P, srv, kind, d, k, shift, bits)
            tickets.append(t)
            outs.append(out)
            if i == 20:  # the failing job sits in the middle of the stream
                bad = srv.submit(P.OP_HIST, bad_keys, P.KEY_U32, shift=0, bits=13, out=bad_out)
        srv.drain()
        st = srv.stats()
        with pytest.raises(P.PAError) as e:
            srv.wait(bad)
        assert e.value.status == 1  # PA_ERR_ARG, with the histogram's own message
        assert "bits" in str(e.value)
        for t in tickets:
            assert srv.wait(t) >= 0
    finally:
        srv.close()
    print(f"admission stats: {st}", flush=True)
    assert st["completed"] == len(mix) and st["failed"] == 1 and st["pending"] == 0
    assert st["peak_bytes_in_flight"] <= budget
    assert st["peak_bytes_in_flight"] >= max(_charge(P, kind, n, bits) for kind, n, _, _, bits in mix)
    assert st["admission_waits"] > 0  # units waited for units in flight
    assert st["oversize_units"] == 0
    assert st["coalesced_jobs"] >= 2
    for (kind, n, k, shift, bits), src, d, out in zip(mix, srcs, dev, outs):
        alone = _solo(P, kind, src, k, shift, bits)
        got = d if kind.startswith("sort") else out
        assert torch.equal(got, alone), (kind, n, k, shift, bits)
        if kind.startswith("sort"):
            assert torch.equal(_as_i64_order(got, kind), torch.sort(_as_i64_order(src, kind)).values), (kind, n)
    torch.cuda.synchronize()


@pytest.mark.parametrize("kind,n,bits", [("sort32", 24 * M, 0), ("sort32", 3000, 0), ("sort64", 12 * M, 0),
                                         ("select32", 32 * M, 0), ("select64", 16 * M, 0), ("hist32", 16 * M, 12),
                                         ("hist64", 16 * M, 5)])
def test_charge_equals_allocation(pa, kind, n, bits):
    """A fresh one-stream server holds exactly what pa_job_scratch_bytes charged."""
    import torch
    P = pa
    d = _keys(P, kind, n, 5)
    torch.cuda.synchronize()
    srv = P.Server(budget_bytes=1 << 30, streams=1)
    try:
        t, _ = _submit(P, srv, kind, d, n // 2, 3, bits)
        srv.wait(t)
        st = srv.stats()
    finally:
        srv.close()
    assert st["bytes_held"] == _charge(P, kind, n, bits)
    assert st["peak_bytes_in_flight"] == st["bytes_held"]


def test_oversize_unit_runs_alone(pa):
    """Budget 32 MiB: a 16M-key sort (about 82 MiB) runs by itself, after the
    other streams' blocks are released, and the jobs around it still run."""
    import torch
    P = pa
    budget = 32 * M
    mix = [("sort32", 1 * M), ("sort32", 2 * M), ("sort32", 16 * M), ("sort32", 1 * M), ("sort64", 1 * M),
           ("sort32", 3 * M)]
    big = _charge(P, "sort32", 16 * M, 0)
    assert big > budget
    srcs = [_keys(P, kind, n, 40 + i) for i, (kind, n) in enumerate(mix)]
    dev = [s.clone() for s in srcs]
    torch.cuda.synchronize()
    srv = P.Server(budget_bytes=budget, streams=2, coalesce_max_keys=1 << 10)
    try:
        tickets = [_submit(P, srv, kind, d, 0, 0, 0)[0] for (kind, n), d in zip(mix, dev)]
        for t in tickets:
            srv.wait(t)
        st = srv.stats()
    finally:
        srv.close()
    print(f"oversize stats: {st}", flush=True)
    assert st["completed"] == len(mix) and st["failed"] == 0
    assert st["oversize_units"] == 1
    assert st["peak_bytes_in_flight"] == big  # the big block, and nothing beside it
    assert st["scratch_releases"] > 0
    for (kind, n), src, d in zip(mix, srcs, dev):
        assert torch.equal(d, _solo(P, kind, src, 0, 0, 0)), (kind, n)


def test_burst_drains_under_small_budget(pa):
    """300 jobs submitted at once under a 48 MiB budget all finish, drain
    returns with nothing pending or in flight, and the held scratch stays
    within the budget."""
    import torch
    P = pa
    budget = 48 * M
    r = P.gen_u64(77, 300)
    sizes = [int(1000 + r[i] % (2 * M)) for i in range(300)]
    srcs = [_keys(P, "sort32", n, 900 + i) for i, n in enumerate(sizes)]
    dev = [s.clone() for s in srcs]
    torch.cuda.synchronize()
    srv = P.Server(budget_bytes=budget, streams=4, coalesce_max_keys=1 << 18)
    try:
        tickets = [srv.submit(P.OP_SORT, d) for d in dev]
        srv.drain()
        st = srv.stats()
        for t in tickets:
            srv.wait(t)
    finally:
        srv.close()
    assert st["completed"] == 300 and st["pending"] == 0
    assert st["peak_bytes_in_flight"] <= budget
    for n, src, d in zip(sizes, srcs, dev):
        assert torch.equal(_as_i64_order(d, "sort32"), torch.sort(_as_i64_order(src, "sort32")).values), n
