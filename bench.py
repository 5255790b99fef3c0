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

# The plain-iNFA serving baseline runs four lanes on streams of their own.  HIP gives
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
    step_ms, kern_ms, engine_ms = [], [], []
    E.Hip.rt().hipDeviceSynchronize()
    for _ in range(steps):
        t0 = time.perf_counter()
        s.scan_device(d_buf.ptr, d_off.ptr, n, kind, mode, use_graph=True)
        res = s.collect(n)  # synchronises the engine stream
        step_ms.append((time.perf_counter() - t0) * 1e3)
        t = s.last_times()
        kern_ms.append(t["sequence"])
        engine_ms.append(t["engine_kernel"])
    E.Hip.rt().hipDeviceSynchronize()
    time_mode.engine_ms = engine_ms
    return step_ms, kern_ms, res


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(path, inputs, budget_s=12.0, threads=1):
    """The oracle on a bounded prefix of the same batch (strings split
    across `threads` threads).  One thread runs pinned to one core of the
    process's affinity set; more threads run on the whole set."""
    import oracle

    allowed = sorted(os.sched_getaffinity(0))
    if threads == 1:
        os.sched_setaffinity(0, {allowed[0]})
    try:
        return _cpu_baseline(oracle, path, inputs, budget_s, threads, allowed)
    finally:
        os.sched_setaffinity(0, set(allowed))


def _cpu_baseline(oracle, path, inputs, budget_s, threads, allowed):
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
    what = (f"single-threaded oracle pinned to core {allowed[0]}" if threads == 1
            else f"oracle on {threads} threads over the {len(allowed)} cores of the affinity set")
    return {"value": round(nbytes / dt / 1e9, 6), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"first {want} of {n} strings ({nbytes} B), {what}, {dt:.1f} s",
            "cpu_model": cpu_model(), "cores_available": len(allowed), "os_cpu_count": os.cpu_count()}


def density_histogram(s, E, d_buf, d_off, n, S):
    """Active states per scan step as a share of all states, on a sample of
    the steps (one in 16, timed by the engine's logging mode): the share of
    steps per density bucket, counting the always-active states."""
    from admatch_amd import selector
    rec = selector.collect(s, d_buf, d_off, n, cap=1 << 22, stride=16)
    act = np.rint(np.exp2(rec[:, 0].astype(np.float64)) - 1)  # total active per logged step
    frac = act / S
    edges = [0, 0.001, 0.003, 0.01, 0.03, 0.10, 1.01]
    names = ["<0.1%", "0.1-0.3%", "0.3-1%", "1-3%", "3-10%", ">=10%"]
    h, _ = np.histogram(frac, bins=edges)
    return {"sampled_steps": int(act.size), "share_of_steps": dict(zip(names, (h / max(1, act.size)).round(4).tolist())),
            "max_active": int(act.max()) if act.size else 0, "max_share": round(float(frac.max()), 4) if act.size else 0,
            "p50_active": float(np.percentile(act, 50)), "p99_active": float(np.percentile(act, 99))}


def pcie_inclusive(s, w, reps=2):
    """BatchedScan from host strings (pinned staging, host-to-device copy,
    scan, result copy and rule mapping), the reference's per-call path."""
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        s.batched_scan_packed(w.inputs.buf, w.inputs.offsets)
        best = min(best, time.perf_counter() - t0)
    nbytes = int(w.inputs.offsets[-1])
    return {"gbs": round(nbytes / best / 1e9, 4), "ms": round(best * 1e3, 2),
            "what": "ita_batched_scan from host memory, H2D and D2H included (best of %d)" % reps}


def other_configs(E, synth, steps):
    """Input GB/s of the small (64 rules, 1,024 strings) and medium (1,000
    rules, 16,384 strings) configurations and of the sparse large rule set,
    inputs resident: plain iNFA (best of 256/512 threads), AS and learned."""
    from admatch_amd import autotune, selector
    out = {}
    for name in ("small", "medium", "large_sparse"):
        w = synth.make_workload(name)
        path = os.path.join(tempfile.mkdtemp(), f"{name}.nfa")
        w.ruleset.write(path)
        s = E.Scratch(E.AD_MATCHER, path).alloc_scratch()
        autotune.apply(s)
        if selector.load_policy(s.info()["hash"]) is None:
            # no table stored for this rule set: compile the network into one
            # on a sample of its traffic drawn apart from the measured strings
            cal = synth.make_workload(name, n=1024, seed_offset=1)
            selector.calibrate(s, cal.inputs.buf, cal.inputs.offsets)
        else:
            selector.apply(s)
        n, nbytes = w.inputs.n, int(w.inputs.offsets[-1])
        d_buf, d_off = E.DeviceArray(w.inputs.buf), E.DeviceArray(w.inputs.offsets)
        r = {"workload": f"{w.ruleset.state_count} states, {n} strings, {nbytes} B", "gbs": {}}
        ref = None
        plain_best = None
        wpb, cap, _, store = s.variant()
        for bs in (256, 512):
            s.set_variant(wpb, cap, bs, store)
            _, kern_ms, res = time_mode(s, E, d_buf, d_off, n, E.KIND_PLAIN, E.MODE_INFA, max(2, steps // 2), 1)
            ref = E.unpack(*res)
            g = nbytes / (float(np.median(kern_ms)) * 1e-3) / 1e9
            plain_best = g if plain_best is None else max(plain_best, g)
        r["gbs"]["plain_infa"] = round(plain_best, 4)
        for mname, mode in (("as", E.MODE_AS), ("learned", E.MODE_LEARNED)):
            _, kern_ms, res = time_mode(s, E, d_buf, d_off, n, E.KIND_ENGINE, mode, steps, 1)
            if E.unpack(*res) != ref:
                raise SystemExit(f"{name}: {mname} disagrees with plain iNFA")
            r["gbs"][mname] = round(nbytes / (float(np.median(kern_ms)) * 1e-3) / 1e9, 4)
            r.setdefault("method_steps", {})[mname] = s.last_stats()[1]
        s.close()
        out[name] = r
        log(f"{name}: {r}")
    return out


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
        # the network compiled into a table on the fixture's other requests
        from admatch_amd import selector
        cb, co = E.pack_strings([z["buf"][z["offsets"][i]:z["offsets"][i + 1]].tobytes()
                                 for i in range(1, z["offsets"].size - 1)])
        selector.calibrate(s, cb, co)
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
        from admatch_amd import selector
        cal = synth.make_workload("medium", n=1024, seed_offset=1)
        selector.calibrate(s, cal.inputs.buf, cal.inputs.offsets)
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
    ids = list(range(n))
    chunks = [[x[c[k]:c[k + 1]] for x, c in zip(strs, cuts)] for k in range(parts)]
    # warm up both paths at full size (buffers, captured launches), untimed
    s.batched_scan(strs)
    st = E.Streams(s, n)
    for k in range(parts):
        st.scan(chunks[k], ids)
    st.close()
    t0 = time.perf_counter()
    whole = s.batched_scan(strs)
    t_whole = time.perf_counter() - t0
    st = E.Streams(s, n)
    t0 = time.perf_counter()
    for k in range(parts):
        got = st.scan(chunks[k], ids)
    t_stream = time.perf_counter() - t0
    phases = {k: round(float(v), 3) for k, v in s.last_times().items()}
    if got != whole:
        raise SystemExit("streaming: chunked results differ from whole-string results")
    st.close()
    s.close()
    out = {"workload": f"{n} strings ({nbytes} B) as {n} streams of {parts} chunks, host strings in",
           "gbs_streamed": round(nbytes / t_stream / 1e9, 4), "gbs_whole_batched": round(nbytes / t_whole / 1e9, 4),
           "last_chunk_call_ms": phases}
    log(f"streaming: {out}")
    return out


def serving(s, E, synth, args):
    """Serving on the large rule set: Poisson arrivals, lognormal lengths
    (median 800 B, max 16 KB), args.serve_seconds of arrivals per rate.  The
    learned engine admits requests on the device (ita_serve_ring: one
    persistent launch, a wave takes the next published request as it
    finishes one); plain iNFA runs the same trace through the host-admitted
    lane loop (ita_serve_trace).  The two must return the same rule ids."""
    from admatch_amd import serve

    rates = [float(r) for r in args.serve_rates.split(",")]
    n_max = int(max(rates) * args.serve_seconds)
    t0 = time.time()
    w = synth.make_workload("serve", n=n_max)
    log(f"serving trace: {n_max} requests generated in {time.time() - t0:.1f} s")
    buf, off = w.inputs.buf, w.inputs.offsets
    # a request's own scan time: 200 requests drawn from the trace, each
    # scanned alone (one launch each, engine-kernel hipEvents); the median
    # over them is what the median latency compares with
    lens = np.diff(off)
    pick = np.random.default_rng(5).choice(min(n_max, 20000), size=200, replace=False)
    own = []
    for i in pick:
        b1, o1 = E.pack_strings([w.inputs.string(int(i))])
        d1, do1 = E.DeviceArray(b1), E.DeviceArray(o1)
        s.scan_device(d1.ptr, do1.ptr, 1, E.KIND_ENGINE, E.MODE_LEARNED)
        s.collect(1)
        own.append(s.last_times()["engine_kernel"] * 1e3)
    own = np.array(own)
    own_us = float(np.median(own))
    out = {"workload": f"{n_max} requests drawn, lognormal median 800 B, max 16 KB, "
                       f"{args.serve_seconds:g} s of arrivals per rate",
           "own_scan": {"sample": "200 requests of the trace, each scanned alone (engine kernel)",
                        "median_us": round(own_us, 1), "mean_us": round(float(own.mean()), 1),
                        "median_bytes": int(np.median(lens[pick]))}, "rates": {}}
    for rate in rates:
        n = int(rate * args.serve_seconds)
        arr = serve.poisson_arrivals(n, rate, seed=int(rate))
        sub = off[: n + 1]
        lat, res, info = serve.serve_ring(s, buf, sub, arr, mode=E.MODE_LEARNED)
        ring = serve.summarize_ring(lat, info, arr)
        ring["p50_over_own_scan"] = round(ring["p50_us"] / own_us, 3)
        lat2, res2, info2 = serve.serve_trace(s, buf, sub, arr, kind=E.KIND_PLAIN, mode=E.MODE_INFA, lanes=4)
        if res2 != res:
            raise SystemExit(f"serving at {rate:.0f} req/s: plain iNFA and the learned engine disagree")
        key = str(int(rate))
        out["rates"][key] = {"rate": rate, "requests": n, "learned_device_admission": ring,
                             "plain_infa_lanes": serve.summarize(lat2, info2, n)}
        log(f"serving {rate:.0f} req/s ({n} requests): learned p50 {ring['p50_us']:.0f} us p99 {ring['p99_us']:.0f} us "
            f"sustained {ring['sustained_req_per_s']:.0f} req/s; plain p50 "
            f"{out['rates'][key]['plain_infa_lanes']['p50_us']:.0f} us")
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
    ap.add_argument("--serve-seconds", type=float, default=1.0,
                    help="seconds of arrivals per rate for the serving measurement (0: skip)")
    ap.add_argument("--serve-rates", default="50000,200000,800000")
    ap.add_argument("--no-streaming", dest="streaming", action="store_false",
                    help="skip the streaming measurement (16,384 strings as 4-chunk streams)")
    ap.add_argument("--no-frontend", dest="frontend", action="store_false",
                    help="skip the regex front-end workload (1,000 compiled patterns x 16,384 strings)")
    ap.add_argument("--no-readme", dest="readme", action="store_false",
                    help="skip the reference README workload (CRS group 6 x bing.txt x 1024)")
    ap.add_argument("--no-extras", dest="extras", action="store_false",
                    help="skip the density histogram, the host-strings rate and the other configurations")
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

    # BatchedScan through this handle (the host-strings rate) follows its
    # flags: the learned selector
    s = E.Scratch(E.AD_MATCHER | E.LEARNED_SELECTOR, path).alloc_scratch()
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
            "kernel_ms": float(np.mean(kern_ms)), "engine_kernel_ms": float(np.mean(time_mode.engine_ms)),
            "p50_step_ms": float(np.percentile(step_ms, 50)),
            "p99_step_ms": float(np.percentile(step_ms, 99)), "method_steps": msteps, "steps": steps,
        }
        log(f"  {name}: {ms:.3f} ms/step, {results[name]['gbs']:.3f} GB/s, kernel {results[name]['kernel_ms']:.3f} ms "
            f"(engine {results[name]['engine_kernel_ms']:.3f}), method steps {msteps}")

    head = "learned" if "learned" in results else list(results)[-1]
    hr = results[head]
    nbytes_alg = total_bytes + 8 * n
    # the dominant kernel's own duration: hipEvents around the engine kernel
    # inside the replayed graph, on its stream, averaged over the timed steps
    kernel_s = (hr["engine_kernel_ms"] or hr["kernel_ms"]) * 1e-3
    achieved = nbytes_alg / kernel_s / 1e9
    traffic = None
    for kname in ("engine_kernel", "engine_h_kernel"):
        pmc = os.path.join(ROOT, "profiles", f"{kname}_{args.config}_pmc.json")
        if os.path.exists(pmc):
            with open(pmc) as f:
                traffic = json.load(f).get("hbm_bytes_per_launch")
            break
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
                      "kernel_ms": round(v["kernel_ms"], 4), "engine_kernel_ms": round(v["engine_kernel_ms"], 4),
                      "method_steps": v["method_steps"]}
                  for k, v in results.items()},
        "speedup_vs_plain_infa": round(hr["gbs"] / results["plain"]["gbs"], 3) if "plain" in results else None,
        "amortized_us_per_request": round(hr["ms_per_step"] * 1e3 / n, 4),
        "p50_us_per_request": None,
        "p99_us_per_request": None,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 4), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
                     "algorithmic_bytes": nbytes_alg, "kernel_ms": round(kernel_s * 1e3, 4)},
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
    if args.serve_seconds > 0 and args.config == "large":
        line["serving"] = serving(s, E, synth, args)
        mid = line["serving"]["rates"][str(int(float(args.serve_rates.split(",")[len(args.serve_rates.split(",")) // 2])))]
        line["p50_us_per_request"] = round(mid["learned_device_admission"]["p50_us"], 2)
        line["p99_us_per_request"] = round(mid["learned_device_admission"]["p99_us"], 2)
        line["latency_source"] = (f"device-admitted serving at {mid['rate']:.0f} req/s "
                                  "(per-request completion seen by the host - arrival)")
    if args.extras and args.config == "large":
        log("density histogram, host-strings rate, other configurations")
        line["active_density"] = density_histogram(s, E, d_buf, d_off, n, w.ruleset.state_count)
        line["pcie_inclusive"] = pcie_inclusive(s, w)
        line["other_configs"] = other_configs(E, synth, max(3, args.steps // 2))
    if not args.no_cpu:
        log("cpu baseline (oracle, 1 thread)")
        line["cpu_baseline"] = cpu_baseline(path, w.inputs)
        threads = len(os.sched_getaffinity(0))  # every core the process may run on
        log(f"cpu baseline (oracle, {threads} threads)")
        line["cpu_baseline_all_cores"] = cpu_baseline(path, w.inputs, budget_s=6.0, threads=threads)
    s.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
