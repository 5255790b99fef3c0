"""Writes the committed test fixtures under tests/golden/.

Run in the build container, where the reference checkout is mounted:

    python tests/golden/make_fixtures.py /root/reference

Outputs (all data, no reference source):
  crs_g{0,6,7,9}.nfa.gz   CRS rule groups of script/run.sh:4-84 merged with
                          admatch_amd.merge (the rules of script/script.py);
                          g6 is the README workload's group, g7 has negative
                          TKO offsets, g9 has word-boundary chains (rule42/47)
  crs_rule1.nfa           data/crs/rule1.nfa as shipped
  http_inputs.npz         request strings: data/bing.txt and
                          data/large_http_request.txt line-joined the way
                          tools/batch.cpp:14-30 does, plus the first 512
                          lines of data/large_request.txt
  crs_expected.json       oracle rule ids for every (group, string)
  synth_small.nfa.gz      the 64-rule synthetic set (seed 1)
  synth_small_inputs.npz  256 of its inputs (seed 101) with planted matches
  synth_small_expected.json  rule ids from Python re.search on each rule's
                          regex source (independent of the oracle)
  ref_nfa_fixtures.tar.gz every per-rule file of data/crs and data/snort as
                          shipped (parser tests, anchored-semantics scans)
"""
from __future__ import annotations

import gzip
import json
import os
import re
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from admatch_amd import merge, synth  # noqa: E402
import oracle  # noqa: E402

GROUPS = (0, 6, 7, 9)


def run_sh_groups(ref: str):
    text = open(os.path.join(ref, "script", "run.sh")).read()
    lists = re.findall(r"for i in \{([0-9,]+)\}", text)
    return [[int(x) for x in l.split(",")] for l in lists]


def batch_join(path: str) -> bytes:
    # tools/batch.cpp:14-30: getline() each line and concatenate, no separator
    data = open(path, "rb").read()
    return b"".join(data.split(b"\n"))


def main(ref: str):
    groups = run_sh_groups(ref)
    crs = os.path.join(ref, "data", "crs")
    texts = {}
    for g in GROUPS:
        paths = [os.path.join(crs, f"rule{i}.nfa") for i in groups[g]]
        paths = [p for p in paths if os.path.exists(p)]
        texts[g] = merge.merge_files(paths)
        with gzip.open(os.path.join(HERE, f"crs_g{g}.nfa.gz"), "wt") as f:
            f.write(texts[g])
    with open(os.path.join(crs, "rule1.nfa"), "rb") as fi, open(os.path.join(HERE, "crs_rule1.nfa"), "wb") as fo:
        fo.write(fi.read())

    strings = [batch_join(os.path.join(ref, "data", "bing.txt")),
               batch_join(os.path.join(ref, "data", "large_http_request.txt"))]
    lines = open(os.path.join(ref, "data", "large_request.txt"), "rb").read().split(b"\n")
    strings += [l for l in lines if l][:512]
    buf, off = synth_pack(strings)
    np.savez_compressed(os.path.join(HERE, "http_inputs.npz"), buf=buf, offsets=off)

    expected = {}
    tmp = os.path.join(HERE, "_tmp.nfa")
    for g in GROUPS:
        with open(tmp, "w") as f:
            f.write(texts[g])
        o = oracle.OracleNFA(tmp)
        expected[f"g{g}"] = o.batch_scan(buf, off, nthreads=8)
    os.remove(tmp)
    with open(os.path.join(HERE, "crs_expected.json"), "w") as f:
        json.dump(expected, f, separators=(",", ":"))

    w = synth.make_workload("small")
    with gzip.open(os.path.join(HERE, "synth_small.nfa.gz"), "wt") as f:
        f.write(w.ruleset.nfa_text)
    planted = [int(i) for i in np.nonzero(w.inputs.planted)[0]]
    others = [i for i in range(w.inputs.n) if not w.inputs.planted[i]]
    idx = sorted(planted + others[: 256 - len(planted)])
    strs = [w.inputs.string(i) for i in idx]
    b2, o2 = synth_pack(strs)
    np.savez_compressed(os.path.join(HERE, "synth_small_inputs.npz"), buf=b2, offsets=o2)
    pats = w.ruleset.patterns()
    exp = [[k + 1 for k, p in enumerate(pats) if p.search(s)] for s in strs]
    with open(os.path.join(HERE, "synth_small_expected.json"), "w") as f:
        json.dump({"rules": len(pats), "expected": exp}, f, separators=(",", ":"))

    # reproducible archive of the per-rule files (fixed owner and mtime)
    subprocess.run(["tar", "czf", os.path.join(HERE, "ref_nfa_fixtures.tar.gz"), "--sort=name", "--owner=0",
                    "--group=0", "--numeric-owner", "--mtime=2020-01-01", "-C", os.path.join(ref, "data"), "crs",
                    "snort"], check=True)


def synth_pack(strings):
    lens = np.array([len(s) for s in strings], dtype=np.int64)
    off = np.zeros(len(strings) + 1, dtype=np.int64)
    off[1:] = np.cumsum(lens)
    return np.frombuffer(b"".join(strings), dtype=np.uint8).copy(), off


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
