"""Pins the CPU oracle: known answers, brute force on tiny NFAs, Python
re.search on the synthetic rules, and the reference fixture statistics."""
import glob
import os

import numpy as np
import pytest

from conftest import golden_inputs, golden_json, golden_nfa, ref_fixture_files
from nfa_gen import brute_force, burst_nfa, burst_strings, random_nfa, shift_nfa

import oracle  # checker under test

REF = "/root/reference"


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_rule1_known_answers(gold_dir):
    # data/crs/rule1.nfa spells (text|application)/xml, unanchored, with a
    # persistent accept state 15 for rule 1
    o = oracle.OracleNFA(golden_nfa("crs_rule1.nfa", gold_dir))
    assert (o.S, o.n_initial, o.n_persistent, o.n_accept_map, o.n_wb) == (20, 1, 2, 1, 0)
    assert o.scan(b"x application/xml y") == [1]
    assert o.scan(b"text/xml") == [1]
    assert o.scan(b"text/xm") == []
    assert o.scan(b"") == []
    assert o.scan(b"TEXT/XML") == []
    assert o.scan(b"application/xhtml+xml") == []
    assert o.scan(b"text/xml" + b"\x00" * 10) == [1]  # accept state persists
    fin = o.final_states(b"text/xml")
    bits = np.unpackbits(fin.view(np.uint8), bitorder="little")
    assert bits[0] == 1 and bits[15] == 1  # start and accept states active


def test_bing_request_on_rule1(gold_dir):
    # tools/batch.cpp line-joins data/bing.txt: 593 bytes, no text/xml in it
    buf, off = golden_inputs("http_inputs.npz")
    assert off[1] - off[0] == 593
    o = oracle.OracleNFA(golden_nfa("crs_rule1.nfa", gold_dir))
    assert o.scan(buf[off[0]:off[1]].tobytes()) == []


@pytest.mark.parametrize("seed", range(40))
def test_brute_force_random_nfas(tmp_path, seed):
    text = random_nfa(seed)
    o = oracle.OracleNFA(write(tmp_path, "r.nfa", text))
    rng = np.random.default_rng(1000 + seed)
    for k in range(25):
        n = int(rng.integers(0, 40))
        data = bytes(rng.choice(list(b"ab_ Z\x00"), size=n))
        assert o.scan(data) == brute_force(text, data), (seed, data)


@pytest.mark.parametrize("seed", range(4))
def test_brute_force_shift_nfas(tmp_path, seed):
    text = shift_nfa(seed, S=300)
    o = oracle.OracleNFA(write(tmp_path, "s.nfa", text))
    rng = np.random.default_rng(seed)
    for _ in range(10):
        data = bytes(rng.choice(list(b"abcx"), size=int(rng.integers(0, 60))))
        assert o.scan(data) == brute_force(text, data)


@pytest.mark.parametrize("seed", range(2))
def test_brute_force_burst_nfas(tmp_path, seed):
    # active sets of up to ~2,000 states that collapse at non-letters
    text = burst_nfa(seed, S=1500)
    o = oracle.OracleNFA(write(tmp_path, "b.nfa", text))
    for data in burst_strings(seed, 6, max_len=200):
        assert o.scan(data) == brute_force(text, data)


def test_word_boundary_chain_order(tmp_path):
    # The pass is in place and in (src,dst) order: 1->2->3 fires fully
    # (edge (1,2) precedes (2,3)); 5->4->0 does not reach 0 in one pass ((4,0) precedes (5,4)).
    text = "\n".join(["6", "1 : initial", "5 : initial", "1 -> 2 : 256", "2 -> 3 : 256", "5 -> 4 : 256",
                      "4 -> 0 : 256", "3 : accepting 7", "0 : accepting 9", "4 : accepting 8"]) + "\n"
    o = oracle.OracleNFA(write(tmp_path, "wb.nfa", text))
    # head boundary before a word char, no steps keep states: check via final states
    assert o.scan(b"") == []
    fin = np.unpackbits(o.final_states(b"").view(np.uint8), bitorder="little")[:6]
    assert fin.tolist() == [0, 1, 0, 0, 0, 1]
    # with persistent states the chain result survives one step
    text2 = text.replace("6\n", "6\n" + "".join(f"{s} -> {s} : 0|255\n" for s in range(6)), 1)
    o2 = oracle.OracleNFA(write(tmp_path, "wb2.nfa", text2))
    # "a": the head pass sets 2, 3 (chain in order) and 4, but not 0, since
    # (4,0) precedes (5,4); the end-of-string pass then reaches 0 from 4
    assert o2.scan(b"a") == [7, 8, 9]
    assert o2.scan(b" ") == []  # no boundary anywhere
    assert o2.scan(b"a") == brute_force(text2, b"a")


def test_synthetic_fixture_vs_re(gold_dir):
    o = oracle.OracleNFA(golden_nfa("synth_small.nfa.gz", gold_dir))
    buf, off = golden_inputs("synth_small_inputs.npz")
    exp = golden_json("synth_small_expected.json")["expected"]
    assert o.batch_scan(buf, off, nthreads=4) == exp
    assert sum(1 for e in exp if e) >= 5


@pytest.mark.parametrize("cfg", ["small", "medium"])
def test_synthetic_fresh_vs_re(workloads, cfg):
    w, path = workloads(cfg, 200)
    o = oracle.OracleNFA(path)
    got = o.batch_scan(w.inputs.buf, w.inputs.offsets, nthreads=8)
    pats = w.ruleset.patterns()
    for i in range(w.inputs.n):
        s = w.inputs.string(i)
        exp = [k + 1 for k, p in enumerate(pats) if p.search(s)]
        assert got[i] == exp, i


def test_crs_expected_fixture(gold_dir):
    buf, off = golden_inputs("http_inputs.npz")
    exp = golden_json("crs_expected.json")
    for g in ("g0", "g6", "g7", "g9"):
        o = oracle.OracleNFA(golden_nfa(f"crs_{g}.nfa.gz", gold_dir))
        assert o.batch_scan(buf, off, nthreads=8) == exp[g], g


def test_threads_do_not_change_results(workloads):
    w, path = workloads("small", 300)
    o = oracle.OracleNFA(path)
    a = o.batch_scan(w.inputs.buf, w.inputs.offsets, nthreads=1)
    b = o.batch_scan(w.inputs.buf, w.inputs.offsets, nthreads=7)
    assert a == b


def test_fixture_statistics(gold_dir):
    # per-set totals of the raw per-rule files (data/crs, data/snort), as
    # counted by text parsing: 18,266 states, 1,875 word-boundary
    # transitions, 867 persistent and 767 accepting states in crs; 405
    # files, 34,443 states, 538 accepting states and no boundary or
    # persistent state in snort
    def totals(files):
        t = dict(states=0, wb=0, persist=0, accept=0, files=0)
        for f in files:
            o = oracle.OracleNFA(f)
            t["states"] += o.S
            t["wb"] += o.n_wb
            t["persist"] += o.n_persistent
            t["accept"] += o.n_accept_map
            t["files"] += 1
        return t
    crs = totals(ref_fixture_files(gold_dir, "crs"))
    snort = totals(ref_fixture_files(gold_dir, "snort"))
    assert (crs["states"], crs["wb"], crs["persist"], crs["accept"]) == (18266, 1875, 867, 767)
    assert (snort["files"], snort["states"], snort["wb"], snort["persist"], snort["accept"]) == (405, 34443, 0, 0, 538)
