"""Benchmark of the batched scan path on one MI355X.

A "step" is one BatchedScan of the whole synthetic batch, inputs already
resident in HBM: the engine kernel (plus its two small length-binning
kernels) replayed from a hipGraph, then the accepted states copied back and
mapped to sorted rule ids on the host.  The default workload is the large
configuration the learned selector is specified on: 3,000 seed-generated
rules (83,408 NFA states, 30% with '.*' gaps) against 65,536 HTTP-like
strings with lognormal lengths (median 1 KB, max 16 KB, 92 MB in total).

Printed: ONE JSON line with input GB/s of the headline selector, the same
for the plain single-kernel iNFA path and the heuristic selectors, per
request latency percentiles, the roofline of the dominant kernel (HBM
bound: 1 input byte + 8 descriptor bytes per string, over the kernel's
average duration from hipEvents on the engine stream) and the CPU oracle's
single-core rate on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# The serving measurement runs four lanes on streams of their own.  HIP gives
# a process GPU_MAX_HW_QUEUES hardware queues (4 by default) and streams beyond
# that share them: with the engine's own stream and torch's already holding
# queues, lanes would pair up and a short batch would wait behind a long one
# queued on the same hardware queue.  Set before the HIP runtime starts.
os.environ["GPU_MAX_HW_QUEUES"] = str(max(int(os.environ.get("GPU_MAX_HW_QUEUES", "4")), 8))

HBM_PEAK_GBS = 8000.0
SELECTOR_PATH = os.path.join(ROOT, "admatch_amd", "data", "selector_weights.npy")


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def time_mode(s, E, d_buf, d_off, n, kind, mode, steps, warmup):
    """Returns (ms per step list, kernel ms list, results of the last step)."""
    res = None
    for _ in range(warmup):
        s.scan_device(d_buf.ptr, d_off.ptr, n, kind, mode, use_graph=True)
        res = s.collect(n)
    step_ms, kern_ms = [], []
    E.Hip.rt().hipDeviceSynchronize()
    for _ in range(steps):
        t0 = time.perf_counter()
        s.scan_device(d_buf.ptr, d_off.ptr, n, kind, mode, use_graph=True)
        res = s.collect(n)  # synchronises the engine stream
        step_ms.append((time.perf_counter() - t0) * 1e3)
        kern_ms.append(s.last_stats()[0])
    E.Hip.rt().hipDeviceSynchronize()
    return step_ms, kern_ms, res


def cpu_baseline(path, inputs, budget_s=12.0, threads=1):
    """The oracle on a bounded prefix of the same batch (strings split
    across `threads` threads)."""
    import oracle

    o = oracle.OracleNFA(path)
    buf, off = inputs.buf, inputs.offsets
    n = off.size - 1
    # calibrate on a few strings, then size the sample to ~budget_s
    k = min(16 * threads, n)
    t0 = time.perf_counter()
    o.batch_scan(buf, off[: k + 1], nthreads=threads)
    dt = max(time.perf_counter() - t0, 1e-6)
    rate = (off[k] - off[0]) / dt
    want = int(min(n, max(k, budget_s * rate / max(1.0, (off[-1] - off[0]) / n))))
    sub = off[: want + 1]
    t0 = time.perf_counter()
    o.batch_scan(buf, sub, nthreads=threads)
    dt = time.perf_counter() - t0
    nbytes = int(sub[-1] - sub[0])
    what = "single-threaded oracle" if threads == 1 else f"oracle on {threads} threads"
    return {"value": round(nbytes / dt / 1e9, 6), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"first {want} of {n} strings ({nbytes} B), {what}, {dt:.1f} s"}


def readme_workload(E, weights, steps):
    """The reference README's benchmark (README.md:13-22): CRS rule group 6
    (script/run.sh:104) against the data/bing.txt request (593 B after
    tools/batch.cpp:14-30 joins its lines) replicated 1,024 times. Per-request
    time = kernel time / 1024 (device-resident inputs) and wall time / 1024
    (host strings in, rule ids out, PCIe included)."""
    import gzip

    gold = os.path.join(ROOT, "tests", "golden")
    path = os.path.join(tempfile.mkdtemp(), "crs_g6.nfa")
    with gzip.open(os.path.join(gold, "crs_g6.nfa.gz"), "rb") as fi, open(path, "wb") as fo:
        fo.write(fi.read())
    z = np.load(os.path.join(gold, "http_inputs.npz"))
    bing = z["buf"][z["offsets"][0]:z["offsets"][1]].tobytes()
    with open(os.path.join(gold, "crs_expected.json")) as f:
        expected = json.load(f)["g6"][0]
    n = 1024
    buf, off = E.pack_strings([bing] * n)
    # ita_batched_scan follows the handle's flags: the learned selector
    s = E.Scratch(E.AD_MATCHER | E.LEARNED_SELECTOR, path).alloc_scratch()
    if weights is not None:
        s.set_selector(weights)
    d_buf, d_off = E.DeviceArray(buf), E.DeviceArray(off)
    out = {"workload": f"CRS group 6 ({s.info()['states']} states) x data/bing.txt request ({len(bing)} B) x {n}",
           "published_us_per_request": {"iNFA": 7258, "AdMatch": 62, "hardware": "not stated (NVIDIA, README.md:19-20)"},
           "us_per_request": {}}
    for name, kind, mode in (("plain_infa", E.KIND_PLAIN, E.MODE_INFA), ("admatch_rule", E.KIND_ENGINE, E.MODE_ADMATCH),
                             ("learned", E.KIND_ENGINE, E.MODE_LEARNED)):
        _, kern_ms, res = time_mode(s, E, d_buf, d_off, n, kind, mode, steps, 1)
        if any(g != expected for g in E.unpack(*res)):
            raise SystemExit(f"README workload: {name} results differ from the golden fixture")
        out["us_per_request"][name] = {"kernel": round(float(np.median(kern_ms)) * 1e3 / n, 3)}
    # host strings in, rule ids out (the reference's per-string figure)
    walls = []
    for _ in range(steps):
        t0 = time.perf_counter()
        got = s.batched_scan_packed(buf, off)
        walls.append(time.perf_counter() - t0)
    if any(g != expected for g in E.unpack(*got)):
        raise SystemExit("README workload: BatchedScan results differ from the golden fixture")
    out["us_per_request"]["learned"]["wall_incl_pcie"] = round(float(np.median(walls)) * 1e6 / n, 3)
    s.close()
    log(f"README workload: {out['us_per_request']}")
    return out


def frontend_workload(E, synth, weights, steps):
    """Rules written as patterns and compiled by the regex front end: 1,000
    patterns of the medium grammar with word boundaries, scanned over the
    medium workload's 16,384 strings (plain iNFA vs the learned engine)."""
    from admatch_amd import regex

    pats = synth.frontend_patterns(1000, seed=11)
    t0 = time.perf_counter()
    text, states = regex.compile_rules(pats, header="front-end workload")
    compile_s = time.perf_counter() - t0
    path = os.path.join(tempfile.mkdtemp(), "frontend.nfa")
    with open(path, "w") as f:
        f.write(text)
    w = synth.make_workload("medium")
    n, nbytes = w.inputs.n, int(w.inputs.offsets[-1])
    s = E.Scratch(E.AD_MATCHER, path).alloc_scratch()
    info = s.info()
    if weights is not None:
        s.set_selector(weights)
    d_buf, d_off = E.DeviceArray(w.inputs.buf), E.DeviceArray(w.inputs.offsets)
    out = {"workload": f"{len(pats)} patterns ({info['states']} states, {info['wb_transitions']} word-boundary "
                       f"transitions) x {n} strings ({nbytes} B)", "compile_s": round(compile_s, 3), "gbs": {}}
    ref = None
    for name, kind, mode in (("plain_infa", E.KIND_PLAIN, E.MODE_INFA), ("as", E.KIND_ENGINE, E.MODE_AS),
                             ("learned", E.KIND_ENGINE, E.MODE_LEARNED)):
        _, kern_ms, res = time_mode(s, E, d_buf, d_off, n, kind, mode, steps, 1)
        res = E.unpack(*res)
        if ref is None:
            ref = res
        elif res != ref:
            raise SystemExit(f"front-end workload: {name} disagrees with plain iNFA")
        out["gbs"][name] = round(nbytes / (float(np.median(kern_ms)) * 1e-3) / 1e9, 4)
    s.close()
    log(f"front-end workload: {out}")
    return out


def streaming_workload(E, path, weights, w, n=16384, parts=4, tune=None):
    """Streaming scans on the large rule set: the first n strings as n
    streams, each cut into `parts` chunks scanned in `parts` calls (host
    strings in, rule ids out) against one BatchedScan of the whole strings."""
    s = E.Scratch(E.AD_MATCHER | E.LEARNED_SELECTOR, path).alloc_scratch()
    if tune:
        s.set_variant(tune["wpb"], tune["cap"], tune.get("plain_bs", 64), tune.get("store", 0))
    if weights is not None:
        s.set_selector(weights)
    strs = [w.inputs.string(i) for i in range(n)]
    nbytes = sum(len(x) for x in strs)
    cuts = [[len(x) * k // parts for k in range(parts + 1)] for x in strs]
    st = E.Streams(s, n)
    ids = list(range(n))
    s.batched_scan(strs[:64])  # warm up
    t0 = time.perf_counter()
    whole = s.batched_scan(strs)
    t_whole = time.perf_counter() - t0
    t0 = time.perf_counter()
    for k in range(parts):
        got = st.scan([x[c[k]:c[k + 1]] for x, c in zip(strs, cuts)], ids)
    t_stream = time.perf_counter() - t0
    if got != whole:
        raise SystemExit("streaming: chunked results differ from whole-string results")
    st.close()
    s.close()
    out = {"workload": f"{n} strings ({nbytes} B) as {n} streams of {parts} chunks, host strings in",
           "gbs_streamed": round(nbytes / t_stream / 1e9, 4), "gbs_whole_batched": round(nbytes / t_whole / 1e9, 4)}
    log(f"streaming: {out}")
    return out


def serving(s, E, synth, args):
    """Continuous-batching request loop on the large rule set: Poisson
    arrivals, lognormal lengths (median 800 B, max 16 KB)."""
    from admatch_amd import serve

    w = synth.make_workload("serve", n=args.serve_requests)
    out = {"workload": f"{w.inputs.n} requests, {int(w.inputs.offsets[-1])} B, lognormal median 800 B",
           "lanes": 4, "rates": {}}
    for r in args.serve_rates.split(","):
        rate = float(r)
        arr = serve.poisson_arrivals(w.inputs.n, rate, seed=int(rate))
        lat, res, info = serve.serve_trace(s, w.inputs.buf, w.inputs.offsets, arr, mode=E.MODE_LEARNED, lanes=4)
        k = min(2000, w.inputs.n)
        lat2, res2, info2 = serve.serve_trace(s, w.inputs.buf, w.inputs.offsets[: k + 1], arr[:k], kind=E.KIND_PLAIN,
                                              mode=E.MODE_INFA, lanes=4)
        if res2 != res[:k]:
            raise SystemExit("serving: plain iNFA and the learned engine disagree")
        out["rates"][r] = {"rate": rate, "learned": serve.summarize(lat, info, w.inputs.n),
                           "plain_infa": serve.summarize(lat2, info2, k)}
        log(f"serving {rate:.0f} req/s: learned p50 {out['rates'][r]['learned']['p50_us']:.0f} us "
            f"p99 {out['rates'][r]['learned']['p99_us']:.0f} us; plain p50 {out['rates'][r]['plain_infa']['p50_us']:.0f} "
            f"p99 {out['rates'][r]['plain_infa']['p99_us']:.0f} us")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="large", choices=["small", "medium", "large", "large_sparse"])
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--modes", default="plain,admatch,heuristic,as,learned")
    ap.add_argument("--serve-requests", type=int, default=20000,
                    help="requests per arrival rate for the serving measurement (0: skip)")
    ap.add_argument("--serve-rates", default="50000,200000,800000")
    ap.add_argument("--no-streaming", dest="streaming", action="store_false",
                    help="skip the streaming measurement (16,384 strings as 4-chunk streams)")
    ap.add_argument("--no-frontend", dest="frontend", action="store_false",
                    help="skip the regex front-end workload (1,000 compiled patterns x 16,384 strings)")
    ap.add_argument("--no-readme", dest="readme", action="store_false",
                    help="skip the reference README workload (CRS group 6 x bing.txt x 1024)")
    ap.add_argument("--variant", default=None,
                    help="wpb,cap,store[,plain_bs]: launch variant (default: the autotuned one)")
    ap.add_argument("--tune", default=None, metavar="PATH",
                    help="autotune the launch variant on this workload first and store it in PATH")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "strings would shard across GPUs by length-balanced shares with a replica of the rule set each, "
              "and nothing but the host result gather crosses GPUs", file=sys.stderr)
        sys.exit(2)

    from admatch_amd import engine as E
    from admatch_amd import synth

    t0 = time.time()
    w = synth.make_workload(args.config, n=args.n)
    n = w.inputs.n
    total_bytes = int(w.inputs.offsets[-1])
    log(f"workload {args.config}: {w.ruleset.state_count} states, {n} strings, {total_bytes} B "
        f"({time.time() - t0:.1f} s to generate)")
    tmp = tempfile.mkdtemp()
    path = os.path.join(tmp, f"{args.config}.nfa")
    w.ruleset.write(path)

    s = E.Scratch(E.AD_MATCHER, path).alloc_scratch()
    info = s.info()
    tuned = None
    tune_path = os.path.join(ROOT, "admatch_amd", "data", "autotune.json")
    if os.path.exists(tune_path):
        with open(tune_path) as f:
            tuned = json.load(f).get(str(info["hash"]))
    if tuned:
        s.set_variant(tuned["wpb"], tuned["cap"], tuned.get("plain_bs", 64), tuned.get("store", 0))
    from admatch_amd import selector as SEL
    learned_ok = SEL.apply(s)
    d_buf, d_off = E.DeviceArray(w.inputs.buf), E.DeviceArray(w.inputs.offsets)
    if args.tune:
        from admatch_amd import autotune
        # tuned in the mode the headline runs
        best = autotune.tune(s, d_buf, d_off, n, mode=E.MODE_LEARNED if learned_ok else E.MODE_AS, reps=1, log=log)
        autotune.store(info["hash"], best, path=args.tune)
    if args.variant:
        v = [int(x) for x in args.variant.split(",")]
        s.set_variant(v[0], v[1], v[3] if len(v) > 3 else s.variant()[2], v[2])
    log(f"variant (wpb, cap, plain_bs, store) = {s.variant()}")

    modes = {
        "plain": (E.KIND_PLAIN, E.MODE_INFA),
        "admatch": (E.KIND_ENGINE, E.MODE_ADMATCH),
        "heuristic": (E.KIND_ENGINE, E.MODE_HEURISTIC),
        "as": (E.KIND_ENGINE, E.MODE_AS),
        "learned": (E.KIND_ENGINE, E.MODE_LEARNED),
        "infa_engine": (E.KIND_ENGINE, E.MODE_INFA),
        "tko": (E.KIND_ENGINE, E.MODE_TKO),
    }
    results, ref = {}, None
    for name in args.modes.split(","):
        kind, mode = modes[name]
        steps = args.steps if name in ("learned",) else max(2, min(args.steps, 5))
        log(f"mode {name}: warmup {args.warmup}, steps {steps}")
        step_ms, kern_ms, res = time_mode(s, E, d_buf, d_off, n, kind, mode, steps, args.warmup)
        out = E.unpack(*res)
        if ref is None:
            ref = out
        elif out != ref:
            raise SystemExit(f"mode {name} disagrees with {list(results)[0]}")
        _, msteps = s.last_stats()
        ms = float(np.mean(step_ms))
        results[name] = {
            "ms_per_step": ms, "gbs": total_bytes / (ms * 1e-3) / 1e9,
            "kernel_ms": float(np.mean(kern_ms)), "p50_step_ms": float(np.percentile(step_ms, 50)),
            "p99_step_ms": float(np.percentile(step_ms, 99)), "method_steps": msteps, "steps": steps,
        }
        log(f"  {name}: {ms:.3f} ms/step, {results[name]['gbs']:.3f} GB/s, kernel {results[name]['kernel_ms']:.3f} ms, "
            f"method steps {msteps}")

    head = "learned" if "learned" in results else list(results)[-1]
    hr = results[head]
    nbytes_alg = total_bytes + 8 * n
    achieved = nbytes_alg / (hr["kernel_ms"] * 1e-3) / 1e9
    traffic = None
    pmc = os.path.join(ROOT, "profiles", f"engine_kernel_{args.config}_pmc.json")
    if os.path.exists(pmc):
        with open(pmc) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    line = {
        "metric": "input GB/s scanned (learned-selector BatchedScan), p50/p99 us/request",
        "value": round(hr["gbs"], 4),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": hr["steps"],
        "warmup": args.warmup,
        "ms_per_step": round(hr["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded rules and HTTP-like strings)",
        "config": {"workload": f"{args.config}: {w.ruleset.state_count} states from "
                               f"{len(w.ruleset.rules)} rules, {n} strings, {total_bytes} B",
                   "variant": dict(zip(("wpb", "cap", "plain_bs", "store"), s.variant())),
                   "learned_selector": learned_ok},
        "modes": {k: {"gbs": round(v["gbs"], 4), "ms_per_step": round(v["ms_per_step"], 4),
                      "kernel_ms": round(v["kernel_ms"], 4), "method_steps": v["method_steps"]}
                  for k, v in results.items()},
        "speedup_vs_plain_infa": round(hr["gbs"] / results["plain"]["gbs"], 3) if "plain" in results else None,
        "amortized_us_per_request": round(hr["ms_per_step"] * 1e3 / n, 4),
        "p50_us_per_request": None,
        "p99_us_per_request": None,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 4), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic},
    }
    if args.readme and args.config == "large":
        line["reference_readme_workload"] = readme_workload(E, np.load(SELECTOR_PATH) if learned_ok else None,
                                                            max(3, args.steps // 2))
    if args.frontend and args.config == "large":
        line["regex_frontend_workload"] = frontend_workload(E, synth, np.load(SELECTOR_PATH) if learned_ok else None,
                                                            max(3, args.steps // 2))
    if args.streaming and args.config == "large":
        line["streaming_workload"] = streaming_workload(E, path, np.load(SELECTOR_PATH) if learned_ok else None, w,
                                                        tune=tuned)
    if args.serve_requests > 0 and args.config == "large":
        line["serving"] = serving(s, E, synth, args)
        mid = line["serving"]["rates"][args.serve_rates.split(",")[len(args.serve_rates.split(",")) // 2]]
        line["p50_us_per_request"] = round(mid["learned"]["p50_us"], 2)
        line["p99_us_per_request"] = round(mid["learned"]["p99_us"], 2)
        line["latency_source"] = f"serving loop at {mid['rate']:.0f} req/s (per-request completion - arrival)"
    if not args.no_cpu:
        log("cpu baseline (oracle, 1 thread)")
        line["cpu_baseline"] = cpu_baseline(path, w.inputs)
        threads = max(1, min(16, os.cpu_count() or 1))  # the box's CPU share for one GPU
        log(f"cpu baseline (oracle, {threads} threads)")
        line["cpu_baseline_all_cores"] = cpu_baseline(path, w.inputs, budget_s=6.0, threads=threads)
    s.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
