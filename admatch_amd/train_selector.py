"""Trains the learned step selector on the GPU and evaluates it.

    python -m admatch_amd.train_selector [--out admatch_amd/data/selector_weights.npy]

1. training set: the large synthetic rule set (3,000 rules) and 8,192
   training strings (seed 104); the engine runs in logging mode and times
   all three step methods on one sampled step in `stride`;
2. training: Adam on the GPU (ita_selector_train), log2-cycle targets;
3. the policy table the engine reads: per (count bucket, list validity,
   byte) entry, the method of least predicted cost summed over the logged
   training steps in the entry (the network's choice at a synthetic step
   context where no step was logged);
4. evaluation on held-out strings of the large workload (seed 103):
   offline regret of the network, of the policy tables and of the
   heuristic on held-out logs, and
   measured scan time of the learned selector vs the heuristic selectors,
   with bit-identical results required;
5. writes the weights (.npy), the table (selector_policy.npz, with the rule
   set's hash) and a JSON report beside them.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

from . import autotune, engine as E, selector, synth


def log(msg):
    print(f"[train_selector] {msg}", file=sys.stderr, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=selector.WEIGHTS_PATH)
    ap.add_argument("--train-n", type=int, default=8192)
    ap.add_argument("--eval-n", type=int, default=8192)
    ap.add_argument("--stride", type=int, default=16)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--tune", action="store_true", help="autotune the launch variant first")
    ap.add_argument("--autotune-out", default=autotune.CACHE, help="autotune cache to update")
    ap.add_argument("--extra-medium", type=int, default=4096,
                    help="also log this many strings on the medium rule set (0: large only)")
    ap.add_argument("--extra-small", type=int, default=2048,
                    help="also log this many strings on the small rule set (0: none)")
    args = ap.parse_args(argv)

    t0 = time.time()
    tr = synth.make_workload("train", n=args.train_n)
    ev = synth.make_workload("large", n=args.eval_n)
    assert tr.ruleset.nfa_text == ev.ruleset.nfa_text
    d = tempfile.mkdtemp()
    path = os.path.join(d, "large.nfa")
    tr.ruleset.write(path)
    log(f"workloads ready in {time.time() - t0:.1f} s: {tr.ruleset.state_count} states")

    s = E.Scratch(E.AD_MATCHER, path).alloc_scratch()
    W = s.info()["words"]
    dtr = (E.DeviceArray(tr.inputs.buf), E.DeviceArray(tr.inputs.offsets))
    dev = (E.DeviceArray(ev.inputs.buf), E.DeviceArray(ev.inputs.offsets))
    if args.tune:
        best = autotune.tune(s, *dev, ev.inputs.n, mode=E.MODE_AS, log=log)
        autotune.store(s.info()["hash"], best, path=args.autotune_out)
    else:
        autotune.apply(s)
    log(f"variant {s.variant()}")

    t0 = time.time()
    rec = selector.collect(s, *dtr, tr.inputs.n, cap=1 << 21, stride=args.stride)
    ds_large = selector.from_log(rec)
    if args.extra_medium > 0:
        # a second rule set (1,000 rules, W = 356 words) so that the network
        # sees more than one state-vector size
        mw = synth.make_workload("medium", n=args.extra_medium, seed_offset=1000)
        mpath = os.path.join(d, "medium.nfa")
        mw.ruleset.write(mpath)
        ms = E.Scratch(E.AD_MATCHER, mpath).alloc_scratch()
        autotune.apply(ms)
        dm = (E.DeviceArray(mw.inputs.buf), E.DeviceArray(mw.inputs.offsets))
        rec = np.concatenate([rec, selector.collect(ms, *dm, mw.inputs.n, cap=1 << 20, stride=args.stride // 2)])
        ms.close()
    if args.extra_small > 0:
        # and a small one (64 rules, W = 33 words)
        sw = synth.make_workload("small", n=args.extra_small, seed_offset=1000)
        spath = os.path.join(d, "small.nfa")
        sw.ruleset.write(spath)
        ss = E.Scratch(E.AD_MATCHER, spath).alloc_scratch()
        autotune.apply(ss)
        dsm = (E.DeviceArray(sw.inputs.buf), E.DeviceArray(sw.inputs.offsets))
        rec = np.concatenate([rec, selector.collect(ss, *dsm, sw.inputs.n, cap=1 << 20, stride=max(1, args.stride // 4))])
        ss.close()
    ds = selector.from_log(rec)
    log(f"training log: {ds.n} steps in {time.time() - t0:.1f} s; best method counts "
        f"{np.bincount(ds.best(), minlength=3).tolist()}")
    rec_ev = selector.collect(s, *dev, ev.inputs.n, cap=1 << 21, stride=args.stride * 4)
    ds_ev = selector.from_log(rec_ev)

    t0 = time.time()
    w, hist = selector.train(ds, epochs=args.epochs, batch=args.batch, lr=args.lr, seed=args.seed)
    log(f"trained {args.epochs} epochs in {time.time() - t0:.1f} s; loss {hist[0]:.4f} -> {hist[-1]:.4f}")

    allow = ds_ev.x[:, 10] > 0
    pred = selector.gpu_forward(ds_ev.x, w)
    rl = selector.regret(ds_ev, selector.choose(pred, allow))
    rh = selector.regret(ds_ev, selector.heuristic_choice(ds_ev, W))
    # what the engine reads: the network folded into the policy table, at
    # synthetic step contexts and as the least expected cost over the
    # training steps of this rule set
    s.set_selector(w, every=0)
    rt = selector.regret(ds_ev, selector.policy_lookup(s.policy_table(), ds_ev.n_expl, allow, ds_ev.c))
    table = selector.policy_from_log(s, w, ds_large)
    rd = selector.regret(ds_ev, selector.policy_lookup(table, ds_ev.n_expl, allow, ds_ev.c))
    s.set_policy(table)
    log(f"held-out steps: network {rl}, synthetic-context table {rt}, table from the log {rd}, heuristic {rh}")

    # measured scans on held-out strings
    timings, ref = {}, None
    for name, mode in (("heuristic", E.MODE_HEURISTIC), ("admatch", E.MODE_ADMATCH), ("learned", E.MODE_LEARNED),
                       ("as", E.MODE_AS)):
        best_ms, out = float("inf"), None
        for _ in range(3):
            s.scan_device(dev[0].ptr, dev[1].ptr, ev.inputs.n, E.KIND_ENGINE, mode)
            out = s.collect(ev.inputs.n)
            best_ms = min(best_ms, s.last_stats()[0])
        if ref is None:
            ref = out
        elif not (np.array_equal(out[0], ref[0]) and np.array_equal(out[1], ref[1])):
            raise SystemExit(f"{name}: results differ from the heuristic's")
        timings[name] = {"kernel_ms": best_ms, "method_steps": s.last_stats()[1]}
        log(f"  {name}: {best_ms:.3f} ms, steps per method {timings[name]['method_steps']}")

    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    np.save(args.out, w)
    selector.save_policy(table, s.info()["hash"], path=os.path.join(os.path.dirname(args.out), "selector_policy.npz"))
    report = {
        "rule_set": {"states": tr.ruleset.state_count, "rules": len(tr.ruleset.rules), "hash": s.info()["hash"]},
        "train": {"strings": tr.inputs.n, "logged_steps": ds.n, "stride": args.stride, "epochs": args.epochs,
                  "batch": args.batch, "lr": args.lr, "loss_first": float(hist[0]), "loss_last": float(hist[-1]),
                  "best_method_counts": np.bincount(ds.best(), minlength=3).tolist()},
        "heldout_steps": {"n": ds_ev.n, "network": rl, "policy_table_synthetic_context": rt,
                          "policy_table_from_log": rd, "heuristic": rh},
        "heldout_scan": {"strings": ev.inputs.n, "bytes": int(ev.inputs.offsets[-1]), **timings},
        "variant": dict(zip(("wpb", "cap", "plain_bs", "store"), s.variant())),
    }
    with open(os.path.splitext(args.out)[0] + "_report.json", "w") as f:
        json.dump(report, f, indent=1)
    print(json.dumps(report))
    s.close()


if __name__ == "__main__":
    main()
