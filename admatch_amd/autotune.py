"""Launch-variant autotuner for the scan engine.

Sweeps the engine's state storage {LDS bitmaps, hashed lists with HBM
bitmaps}, waves per workgroup {1, 2, 4} and active-list capacity
{128, 256, 512, 1024}, and the plain iNFA kernel's workgroup size {64, ..., 1024},
on a device-resident sample of the workload; keeps the fastest by hipEvent
kernel time (all variants must give identical results).  Results are cached
in admatch_amd/data/autotune.json keyed by the rule set's content hash.
"""
from __future__ import annotations

import json
import os
from typing import Dict

import numpy as np

from . import engine as E

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "autotune.json")


def _time(s: E.Scratch, d_buf, d_off, n: int, kind: int, mode: int, reps: int) -> float:
    s.scan_device(d_buf.ptr, d_off.ptr, n, kind, mode)
    s.collect(n)
    best = float("inf")
    for _ in range(reps):
        s.scan_device(d_buf.ptr, d_off.ptr, n, kind, mode)
        s.collect(n)
        best = min(best, s.last_stats()[0])
    return best


def tune(s: E.Scratch, d_buf, d_off, n: int, mode: int = E.MODE_HEURISTIC, reps: int = 2,
         plain: bool = True, log=print) -> Dict:
    info = s.info()
    results = []
    ref = None
    stores = (0, 1) if mode in (E.MODE_AS, E.MODE_LEARNED) else (0,)
    for store in stores:
        for wpb in (1, 2, 4):
            for cap in (128, 256, 512, 1024):
                try:
                    s.set_variant(wpb, cap, 64, store)
                    ms = _time(s, d_buf, d_off, n, E.KIND_ENGINE, mode, reps)
                except E.AdmatchError as e:  # does not fit the LDS budget
                    log(f"  store={store} wpb={wpb} cap={cap}: skipped ({e})")
                    continue
                out = s.collect(n)
                if ref is None:
                    ref = out
                elif not (np.array_equal(out[0], ref[0]) and np.array_equal(out[1], ref[1])):
                    raise RuntimeError(f"variant store={store} wpb={wpb} cap={cap} changed the results")
                log(f"  store={store} wpb={wpb} cap={cap}: {ms:.3f} ms")
                results.append((ms, wpb, cap, store))
    ms, wpb, cap, store = min(results)
    best = {"wpb": wpb, "cap": cap, "store": store, "engine_ms": ms, "mode": E.MODE_NAMES[mode],
            "states": info["states"],
            "sweep": [{"wpb": w, "cap": c, "store": st, "ms": m} for m, w, c, st in results]}
    plain_bs = 64
    if plain and info["plain_lds_bytes"] <= 160 * 1024:
        pr = []
        for bs in (64, 128, 256, 512, 1024):
            s.set_variant(wpb, cap, bs, store)
            m = _time(s, d_buf, d_off, n, E.KIND_PLAIN, E.MODE_INFA, 1)
            log(f"  plain bs={bs}: {m:.3f} ms")
            pr.append((m, bs))
        plain_bs = min(pr)[1]
        best["plain_sweep"] = [{"bs": b, "ms": m} for m, b in pr]
    best["plain_bs"] = plain_bs
    s.set_variant(wpb, cap, plain_bs, store)
    return best


def load(hash_: int):
    if not os.path.exists(CACHE):
        return None
    with open(CACHE) as f:
        return json.load(f).get(str(hash_))


def store(hash_: int, entry: Dict, path: str = CACHE) -> None:
    data = {}
    if os.path.exists(path):
        with open(path) as f:
            data = json.load(f)
    data[str(hash_)] = entry
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)


def apply(s: E.Scratch) -> bool:
    e = load(s.info()["hash"])
    if not e:
        return False
    s.set_variant(e["wpb"], e["cap"], e.get("plain_bs", 64), e.get("store", 0))
    return True
