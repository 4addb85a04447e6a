"""The continuous-batching server: a mixed stream of sort / select / histogram
jobs submitted while others run must give, for every job, exactly the result
of the same call made alone (and the oracle's)."""
import numpy as np
import pytest

from gpu_util import dev_empty, to_dev, to_host

pytestmark = pytest.mark.gpu


def _jobs(seed, count):
    import parallelalgorithms_amd as P
    r = P.gen_u64(seed, 4 * count)
    jobs = []
    for i in range(count):
        u = int(r[i] % 1000)
        n = int(np.exp(np.log(4096) + (int(r[count + i] >> np.uint64(11)) / float(1 << 53)) * (np.log(4_000_000) - np.log(4096))))
        if u < 700:
            kind = "sort64" if (int(r[2 * count + i]) & 1) else "sort32"
        elif u < 900:
            kind = "select64" if (int(r[2 * count + i]) & 1) else "select"
        else:
            kind = "hist64" if (int(r[2 * count + i]) & 1) else "hist"
        if kind == "sort32" and i % 3 == 0:
            n = int(n % 50_000) + 2  # small ones: coalesced into batches
        keys = P.gen_u64(seed + 7 * i, n) if kind.endswith("64") else P.gen_u32(seed + 7 * i, n)
        k = int(r[3 * count + i] % max(1, n // 10))
        jobs.append((kind, keys, k))
    return jobs


def test_server_mix_equals_alone(pa, orc):
    import parallelalgorithms_amd as P
    jobs = _jobs(1240, 60)
    srv = P.Server(budget_bytes=2 << 30, streams=4, coalesce_max_keys=1 << 16)
    try:
        # Device inputs first, then every job submitted in one burst, so the
        # small sorts are waiting together while the large jobs run (the
        # uploads would otherwise space the submissions out).
        import torch
        dev = [to_dev(keys) for _, keys, _ in jobs]
        torch.cuda.synchronize()
        outs, tickets = [], []
        for (kind, keys, k), d in zip(jobs, dev):
            out = None
            if kind == "select":
                out = dev_empty(1, np.uint32)
                t = srv.submit(P.OP_SELECT, d, P.KEY_U32, k=k, out=out)
            elif kind == "select64":
                out = dev_empty(1, np.uint64)
                t = srv.submit(P.OP_SELECT, d, P.KEY_U64, k=k, out=out)
            elif kind == "hist":
                out = dev_empty(256, np.uint64)
                t = srv.submit(P.OP_HIST, d, P.KEY_U32, shift=24, bits=8, out=out)
            elif kind == "hist64":
                out = dev_empty(256, np.uint64)
                t = srv.submit(P.OP_HIST, d, P.KEY_U64, shift=56, bits=8, out=out)
            else:
                t = srv.submit(P.OP_SORT, d, P.KEY_U64 if kind == "sort64" else P.KEY_U32)
            outs.append(out)
            tickets.append(t)
        lat = [srv.wait(t) for t in tickets]
        st = srv.stats()
    finally:
        srv.close()
    assert st["completed"] == len(jobs) and st["failed"] == 0
    assert st["coalesced_jobs"] >= 2  # small sorts were batched
    assert all(x >= 0 for x in lat)
    for (kind, keys, k), d, out in zip(jobs, dev, outs):
        if kind == "sort32":
            got = to_host(d, np.uint32)
            alone = to_dev(keys)
            pa.sort_u32(alone)
            np.testing.assert_array_equal(got, to_host(alone, np.uint32))
            np.testing.assert_array_equal(got, np.sort(keys))
        elif kind == "sort64":
            got = to_host(d, np.uint64)
            np.testing.assert_array_equal(got, np.sort(keys))
        elif kind == "select":
            expect, _ = orc.select_radix_u32(keys, k)
            assert int(to_host(out, np.uint32)[0]) == expect
        elif kind == "select64":
            assert int(to_host(out, np.uint64)[0]) == int(np.sort(keys)[k])
        elif kind == "hist":
            np.testing.assert_array_equal(to_host(out, np.uint64), orc.hist_one_byte_u32(keys, 0, keys.size - 1, 24))
        else:
            np.testing.assert_array_equal(to_host(out, np.uint64),
                                          np.bincount((keys >> np.uint64(56)).astype(np.int64), minlength=256))


def test_server_rejects_host_memory(pa):
    import parallelalgorithms_amd as P
    srv = P.Server(budget_bytes=1 << 30, streams=1)
    try:
        with pytest.raises(P.PAError):
            srv.submit(P.OP_SORT, np.arange(100, dtype=np.uint32))
    finally:
        srv.close()
