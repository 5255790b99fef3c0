"""GPU: the reference's C++ surface (include/ITA_regex.h) from a C++ program.

admatch_amd/bin/ita_caller is a user program of ita_scratch / ad_scratch,
allocScratch, Scan, BatchedScan and freeScratch.  Its output vectors arrive
pre-filled, so the results must be the union of the pre-filled ids and the
accepted rules, sorted and de-duplicated (src/iNFA_kernel.cu:298-306).  The
two reference command-line tools (bin/batch, bin/sequence) run on inputs too.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

from conftest import ROOT, golden_inputs, golden_json, golden_nfa

pytestmark = pytest.mark.gpu

from admatch_amd import engine as E  # noqa: E402

import oracle  # noqa: E402  (checker only)

BIN = os.path.join(ROOT, "admatch_amd", "bin")
FLAGS = [("iNFA", E.INFA_KERNEL), ("TKO", E.TKO_KERNEL), ("AS", E.AS_KERNEL), ("AD", E.AD_MATCHER),
         ("learned", E.AD_MATCHER | E.LEARNED_SELECTOR)]


def prefill(i, expected):
    # three kinds of pre-filled vectors: empty, ids beyond any rule (plus a
    # duplicate), and the expected ids themselves in reverse with a 0
    if i % 3 == 0:
        return []
    if i % 3 == 1:
        return [999999, 5, 999999]
    return list(reversed(expected)) + [0]


def write_requests(path, strings, pre):
    with open(path, "wb") as f:
        f.write(struct.pack("<I", len(strings)))
        for s, p in zip(strings, pre):
            f.write(struct.pack("<I", len(s)) + s + struct.pack("<I", len(p)))
            f.write(np.asarray(p, dtype="<i4").tobytes())


def run_caller(tmp_path, flags, nfa, strings, pre, single=False):
    req, out = str(tmp_path / "req.bin"), str(tmp_path / "out.txt")
    write_requests(req, strings, pre)
    cmd = [os.path.join(BIN, "ita_caller"), str(flags), nfa, req, out] + (["scan"] if single else [])
    print(f"[case] {' '.join(cmd[1:3])} n={len(strings)} single={single}", flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    with open(out) as f:
        return [[int(x) for x in line.split()] for line in f.read().splitlines()]


@pytest.mark.parametrize("name,flags", FLAGS, ids=[f[0] for f in FLAGS])
@pytest.mark.parametrize("group", ["g6", "g9"])
def test_batched_scan_union_semantics(tmp_path, gold_dir, group, name, flags):
    nfa = golden_nfa(f"crs_{group}.nfa.gz", gold_dir)
    buf, off = golden_inputs("http_inputs.npz")
    exp = golden_json("crs_expected.json")[group]
    strings = [buf[off[i]:off[i + 1]].tobytes() for i in range(off.size - 1)]
    pre = [prefill(i, e) for i, e in enumerate(exp)]
    got = run_caller(tmp_path, flags, nfa, strings, pre)
    want = [sorted(set(p) | set(e)) for p, e in zip(pre, exp)]
    assert got == want


@pytest.mark.parametrize("name,flags", [FLAGS[0], FLAGS[3]], ids=["iNFA", "AD"])
def test_scan_one_string_at_a_time(tmp_path, gold_dir, name, flags):
    nfa = golden_nfa("crs_g9.nfa.gz", gold_dir)
    buf, off = golden_inputs("http_inputs.npz")
    exp = golden_json("crs_expected.json")["g9"]
    idx = list(range(0, off.size - 1, 8))
    strings = [buf[off[i]:off[i + 1]].tobytes() for i in idx]
    assert all(b"\0" not in s for s in strings)  # Scan takes strlen
    pre = [prefill(k, exp[i]) for k, i in enumerate(idx)]
    got = run_caller(tmp_path, flags, nfa, strings, pre, single=True)
    assert got == [sorted(set(p) | set(exp[i])) for p, i in zip(pre, idx)]


def test_batched_scan_many_rules_per_string(tmp_path):
    # 300 rules accept on any string holding an 'a': the ids outgrow the
    # shim's first output buffer (4 per string), whose results are then
    # fetched from the handle instead of scanned again
    R = 300
    lines = [str(R + 1), "0 : initial", "0 -> 0 : 0|255"]
    for r in range(1, R + 1):
        lines += [f"0 -> {r} : 97", f"{r} : accepting {r}", f"{r} -> {r} : 0|255"]
    nfa = str(tmp_path / "many.nfa")
    with open(nfa, "w") as f:
        f.write("\n".join(lines) + "\n")
    rng = np.random.default_rng(5)
    strings = [bytes(rng.choice(list(b"abc "), size=int(rng.integers(0, 60)))) for _ in range(64)]
    exp = oracle.OracleNFA(nfa).batch_scan(*E.pack_strings(strings))
    assert sum(len(e) for e in exp) > 4 * len(strings) + 16
    pre = [prefill(i, e) for i, e in enumerate(exp)]
    for flags in (E.INFA_KERNEL, E.AD_MATCHER):
        got = run_caller(tmp_path, flags, nfa, strings, pre)
        assert got == [sorted(set(p) | set(e)) for p, e in zip(pre, exp)]


def test_sequence_tool(gold_dir):
    # tools/sequence.cpp:11-17: one string, PROFILER_MODE | SHOW_RESULTS
    nfa = golden_nfa("crs_rule1.nfa", gold_dir)
    r = subprocess.run([os.path.join(BIN, "sequence"), nfa, "x application/xml y"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    out = r.stdout
    fin = oracle.OracleNFA(nfa).final_states(b"x application/xml y")
    states = np.nonzero(np.unpackbits(fin.view(np.uint8), bitorder="little"))[0].tolist()
    assert f"Final States : {' '.join(map(str, states))} \n" in out
    assert "Accept Rules : 1 \n" in out
    for label in ("Total:", "Device MemAlloc:", "MemCpy H2D:", "Kernel Exec:", "MemCpy D2H:", "Device MemFree:"):
        assert label in out, label


@pytest.mark.parametrize("kernel", ["iNFA", "AD", "LEARNED"])
def test_batch_tool_readme_workload(tmp_path, gold_dir, kernel):
    # the README workload's command line (script/run.sh:104) with -p and -s,
    # on CRS group 9 and a line of data/large_request.txt that matches rules
    nfa = golden_nfa("crs_g9.nfa.gz", gold_dir)
    buf, off = golden_inputs("http_inputs.npz")
    exp_all = golden_json("crs_expected.json")["g9"]
    k = next(i for i, e in enumerate(exp_all) if e)
    inp = tmp_path / "req.txt"
    inp.write_bytes(buf[off[k]:off[k + 1]].tobytes() + b"\n")
    exp = exp_all[k]
    r = subprocess.run([os.path.join(BIN, "batch"), "-f", nfa, "-i", str(inp), "-n", "8", "-k", kernel, "-p", "-s"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rules = [l for l in r.stdout.splitlines() if l.startswith("Accept Rules :")]
    assert rules == ["Accept Rules : " + "".join(f"{x} " for x in exp)] * 8
    assert r.stdout.count("Final States :") == 8 and "Kernel Exec:" in r.stdout
