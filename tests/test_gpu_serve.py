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
