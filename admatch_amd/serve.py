"""Serving drivers: Poisson request traces replayed through the native loops
(admatch_amd/csrc/serve.cpp).

    python -m admatch_amd.serve --rates 50000,200000,800000 --seconds 1 [--plain]

Requests are strings of the serving workload (the large 3,000-rule set,
lognormal lengths with median 800 B, max 16 KB).  For each arrival rate the
learned-selector engine serves the trace with device-side admission
(ita_serve_ring: one persistent launch, a wave takes the next published
request as it finishes one); with --plain, the plain iNFA kernel serves the
same trace through the host-admitted lane loop (ita_serve_trace).  Reported
per rate: p50/p99 latency and the sustained request and byte rates.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import tempfile

import numpy as np

from . import autotune, engine as E, selector, synth


def _sig():
    L = E.lib()
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    L.ita_serve_trace.restype = i32
    L.ita_serve_trace.argtypes = [vp, i32, i32, i32, i32, i64, i64, i64, vp, vp, vp, i32, vp, vp, vp, i64, vp]
    return L


def poisson_arrivals(n: int, rate: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


def serve_trace(s: E.Scratch, buf: np.ndarray, offsets: np.ndarray, arrival_s: np.ndarray, kind: int = E.KIND_ENGINE,
                mode: int = E.MODE_LEARNED, lanes: int = 4, max_batch: int = 4096, max_batch_bytes: int = 8 << 20,
                state_budget: int = 256 << 20, split_bytes: int = 2048):
    buf = np.ascontiguousarray(buf, dtype=np.uint8)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    arrival_s = np.ascontiguousarray(arrival_s, dtype=np.float64)
    n = offsets.size - 1
    lat = np.zeros(max(n, 1), dtype=np.float64)
    out_off = np.zeros(n + 1, dtype=np.int32)
    stats = np.zeros(6, dtype=np.float64)
    cap = max(4 * n, 1024)
    while True:
        ids = np.zeros(cap, dtype=np.int32)
        rc = _sig().ita_serve_trace(s._h, kind, mode, lanes, max_batch, max_batch_bytes, state_budget, split_bytes,
                                    buf.ctypes.data, offsets.ctypes.data, arrival_s.ctypes.data, n,
                                    lat.ctypes.data, out_off.ctypes.data, ids.ctypes.data, cap, stats.ctypes.data)
        if rc == E.ITA_ERR_CAPACITY and int(out_off[-1]) > cap:
            cap = int(out_off[-1])
            continue
        E._err(rc)
        break
    info = {"wall_s": stats[0], "batches": int(stats[1]), "mean_batch": stats[2],
            "peak_inflight_bytes": int(stats[3]), "peak_lanes_busy": int(stats[4]), "bytes": int(stats[5])}
    return lat[:n], E.unpack(out_off, ids[: out_off[-1]]), info


def serve_ring(s: E.Scratch, buf: np.ndarray, offsets: np.ndarray, arrival_s: np.ndarray,
               mode: int = E.MODE_LEARNED):
    """Device-side serving (ita_serve_ring): one persistent launch whose
    waves take each request as soon as the host publishes its arrival."""
    L = E.lib()
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    L.ita_serve_ring.restype = i32
    L.ita_serve_ring.argtypes = [vp, i32, vp, vp, vp, i32, vp, vp, vp, i64, vp]
    buf = np.ascontiguousarray(buf, dtype=np.uint8)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    arrival_s = np.ascontiguousarray(arrival_s, dtype=np.float64)
    n = offsets.size - 1
    lat = np.zeros(max(n, 1), dtype=np.float64)
    out_off = np.zeros(n + 1, dtype=np.int32)
    stats = np.zeros(6, dtype=np.float64)
    cap = max(4 * n, 1024)
    ids = np.zeros(cap, dtype=np.int32)
    rc = L.ita_serve_ring(s._h, mode, buf.ctypes.data, offsets.ctypes.data, arrival_s.ctypes.data, n,
                          lat.ctypes.data, out_off.ctypes.data, ids.ctypes.data, cap, stats.ctypes.data)
    if rc == E.ITA_ERR_CAPACITY:  # the results are kept by the handle
        ids = np.zeros(int(out_off[-1]), dtype=np.int32)
        L.ita_last_results(s._h, ids.ctypes.data, ids.size)
    else:
        E._err(rc)
    info = {"span_s": stats[0], "rescanned": int(stats[1]), "waves": int(stats[2]),
            "method_steps": [int(x) for x in stats[3:6]], "bytes": int(offsets[-1] - offsets[0])}
    return lat[:n], E.unpack(out_off, ids[: out_off[-1]]), info


def summarize_ring(lat: np.ndarray, info: dict, arrival_s: np.ndarray) -> dict:
    n = lat.size
    span = max(info["span_s"], 1e-9)
    return {"p50_us": float(np.percentile(lat, 50) * 1e6), "p99_us": float(np.percentile(lat, 99) * 1e6),
            "mean_us": float(lat.mean() * 1e6), "offered_req_per_s": float(n / max(arrival_s[-1] - arrival_s[0], 1e-9)),
            "sustained_req_per_s": n / span, "gbs": info["bytes"] / span / 1e9, "waves": info["waves"],
            "method_steps": info["method_steps"], "rescanned": info["rescanned"]}


def summarize(lat: np.ndarray, info: dict, n: int) -> dict:
    return {"p50_us": float(np.percentile(lat, 50) * 1e6), "p99_us": float(np.percentile(lat, 99) * 1e6),
            "mean_us": float(lat.mean() * 1e6), "req_per_s": n / info["wall_s"],
            "gbs": info["bytes"] / info["wall_s"] / 1e9, "batches": info["batches"],
            "mean_batch": round(info["mean_batch"], 1), "peak_lanes_busy": info["peak_lanes_busy"]}


def hw_queues_for(lanes: int) -> None:
    """Give each lane a hardware queue of its own: HIP gives a process
    GPU_MAX_HW_QUEUES queues (4 by default) and streams beyond that share one,
    so a short batch would wait behind a long one queued on the same queue.
    Takes effect only before the HIP runtime starts."""
    want = lanes + 4  # the engine's own stream, the null stream, spare
    os.environ["GPU_MAX_HW_QUEUES"] = str(max(int(os.environ.get("GPU_MAX_HW_QUEUES", "4")), want))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--rates", default="50000,200000,800000")
    ap.add_argument("--seconds", type=float, default=1.0, help="seconds of arrivals per rate")
    ap.add_argument("--plain", action="store_true", help="also run plain iNFA through the lane loop")
    args = ap.parse_args(argv)
    hw_queues_for(4)
    rates = [float(r) for r in args.rates.split(",")]
    w = synth.make_workload("serve", n=int(max(rates) * args.seconds))
    d = tempfile.mkdtemp()
    path = os.path.join(d, "serve.nfa")
    w.ruleset.write(path)
    s = E.Scratch(E.AD_MATCHER | E.LEARNED_SELECTOR, path).alloc_scratch()
    autotune.apply(s)
    learned = selector.apply(s)
    report = {"workload": f"{w.ruleset.state_count} states, {w.inputs.n} requests, "
                          f"{int(w.inputs.offsets[-1])} B", "learned_selector": learned, "rates": {}}
    for rate in rates:
        n = int(rate * args.seconds)
        arr = poisson_arrivals(n, rate, seed=int(rate))
        sub = w.inputs.offsets[: n + 1]
        lat, res, info = serve_ring(s, w.inputs.buf, sub, arr, mode=E.MODE_LEARNED)
        r = {"engine_device_admission": summarize_ring(lat, info, arr)}
        if args.plain:
            lat2, res2, info2 = serve_trace(s, w.inputs.buf, sub, arr, kind=E.KIND_PLAIN, mode=E.MODE_INFA, lanes=4)
            assert res2 == res, "plain iNFA and the engine disagree"
            r["plain_infa_lanes"] = summarize(lat2, info2, n)
        report["rates"][str(int(rate))] = r
        print(f"[serve] rate {rate:.0f}/s {r}", file=sys.stderr, flush=True)
    print(json.dumps(report))
    s.close()


if __name__ == "__main__":
    main()
