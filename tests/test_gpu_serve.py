"""Continuous-batching serving loop: every request's result is the same
alone and in the mix; admission respects the state budget."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from admatch_amd import engine as E  # noqa: E402
from admatch_amd import serve  # noqa: E402

import oracle  # noqa: E402


@pytest.mark.parametrize("rate,lanes,budget,split", [(2e5, 4, 256 << 20, 2048), (2e6, 3, 64 << 10, 0),
                                                     (5e4, 1, 256 << 20, 1024)])
def test_mix_equals_alone(workloads, rate, lanes, budget, split):
    w, path = workloads("medium", 1500)
    buf, off = w.inputs.buf, w.inputs.offsets
    n = off.size - 1
    s = E.Scratch(E.AD_MATCHER | E.HEURISTIC_SELECTOR, path).alloc_scratch()
    arr = serve.poisson_arrivals(n, rate, seed=1)
    lat, res, info = serve.serve_trace(s, buf, off, arr, mode=E.MODE_HEURISTIC, lanes=lanes, max_batch=256,
                                       state_budget=budget, split_bytes=split)
    assert info["batches"] >= 2 and info["peak_inflight_bytes"] <= budget
    assert (lat > 0).all()
    # alone: each request scanned by itself
    alone = [s.batched_scan([w.inputs.string(i)])[0] for i in range(0, n, 7)]
    assert [res[i] for i in range(0, n, 7)] == alone
    # and the whole mix equals the oracle
    assert res == oracle.OracleNFA(path).batch_scan(buf, off, nthreads=8)
    s.close()


def test_overlapping_lanes(workloads):
    # a burst larger than one batch keeps several lanes busy at once
    w, path = workloads("medium", 2000)
    s = E.Scratch(E.AD_MATCHER, path).alloc_scratch()
    arr = np.zeros(w.inputs.n)
    lat, res, info = serve.serve_trace(s, w.inputs.buf, w.inputs.offsets, arr, lanes=4, max_batch=128)
    assert info["peak_lanes_busy"] >= 2
    assert res == oracle.OracleNFA(path).batch_scan(w.inputs.buf, w.inputs.offsets, nthreads=8)
    s.close()


def test_budget_rejects_oversized_request(workloads):
    w, path = workloads("small", 8)
    s = E.Scratch(E.AD_MATCHER, path).alloc_scratch()
    with pytest.raises(E.AdmatchError) as ei:
        serve.serve_trace(s, w.inputs.buf, w.inputs.offsets, np.zeros(w.inputs.n), state_budget=100)
    assert ei.value.rc == -7
    s.close()


# ---------------------------------------------------------------- the large rule set (configuration 5)
@pytest.fixture(scope="module")
def serve_sample(workloads):
    """Requests of the serving workload (the 3,000-rule large set, lognormal
    lengths, median 800 B, max 16 KB): the first 3,000 of a 20,000-request
    draw plus its 64 longest, with the oracle's rule ids."""
    w, path = workloads("serve", 20000)
    lens = np.diff(w.inputs.offsets)
    idx = sorted(set(range(3000)) | set(np.argsort(lens)[-64:].tolist()))
    strs = [w.inputs.string(i) for i in idx]
    buf, off = E.pack_strings(strs)
    exp = oracle.OracleNFA(path).batch_scan(buf, off, nthreads=16)
    assert max(len(x) for x in strs) >= 8192 and sum(1 for e in exp if e) > 0
    return path, strs, buf, off, exp


def _learned_scratch(path, store):
    from admatch_amd import autotune, selector
    s = E.Scratch(E.AD_MATCHER | E.LEARNED_SELECTOR, path).alloc_scratch()
    autotune.apply(s)
    wpb, cap, pbs, _ = s.variant()
    if store == 1:
        wpb, cap = 2, 256
    s.set_variant(wpb, cap, pbs, store)
    assert selector.apply(s)
    return s


@pytest.mark.parametrize("store", [0, 1])
def test_ring_large_rule_set_vs_oracle(serve_sample, store):
    # device-side admission: each request's rule ids equal the oracle's at
    # every rate, and equal the request scanned alone
    path, strs, buf, off, exp = serve_sample
    s = _learned_scratch(path, store)
    n = off.size - 1
    for rate in (5e4, 2e5, 8e5):
        print(f"[case] ring store={store} rate={rate:.0f} n={n}", flush=True)
        arr = serve.poisson_arrivals(n, rate, seed=int(rate))
        lat, res, info = serve.serve_ring(s, buf, off, arr, mode=E.MODE_LEARNED)
        assert res == exp, next(i for i in range(n) if res[i] != exp[i])
        assert (lat > 0).all() and info["waves"] > 0
    alone = [s.batched_scan([strs[i]])[0] for i in range(0, n, 61)]
    assert alone == [exp[i] for i in range(0, n, 61)]
    s.close()


@pytest.mark.parametrize("store", [0, 1])
def test_lanes_large_rule_set_vs_oracle(serve_sample, store):
    # the host-admitted lane loop on the same requests, learned selector,
    # both storages (hashed: one device bitmap buffer per lane)
    path, strs, buf, off, exp = serve_sample
    s = _learned_scratch(path, store)
    n = off.size - 1
    for rate in (5e4, 2e5, 8e5):
        print(f"[case] lanes store={store} rate={rate:.0f} n={n}", flush=True)
        arr = serve.poisson_arrivals(n, rate, seed=int(rate))
        lat, res, info = serve.serve_trace(s, buf, off, arr, mode=E.MODE_LEARNED, lanes=4)
        assert res == exp, next(i for i in range(n) if res[i] != exp[i])
    s.close()


def test_ring_edge_cases(workloads):
    # empty requests, all arriving at once, a single request, an empty trace
    w, path = workloads("small", 300)
    s = E.Scratch(E.AD_MATCHER, path).alloc_scratch()
    strs = [b""] + [w.inputs.string(i) for i in range(299)] + [b""]
    buf, off = E.pack_strings(strs)
    exp = oracle.OracleNFA(path).batch_scan(buf, off, nthreads=8)
    lat, res, _ = serve.serve_ring(s, buf, off, np.zeros(len(strs)), mode=E.MODE_HEURISTIC)
    assert res == exp
    b1, o1 = E.pack_strings(strs[1:2])
    assert serve.serve_ring(s, b1, o1, np.array([0.001]))[1] == exp[1:2]
    assert serve.serve_ring(s, np.zeros(0, np.uint8), np.zeros(1, np.int64), np.zeros(0))[1] == []
    with pytest.raises(E.AdmatchError):
        serve.serve_ring(s, buf, off, np.arange(len(strs))[::-1].astype(np.float64))  # decreasing arrivals
    s.close()
