"""GPU parity: every kernel and selector against the CPU oracle (and the
golden fixtures), through the C ABI.  Bit-exact on sorted rule-id lists."""
import os

import numpy as np
import pytest

from conftest import golden_inputs, golden_json, golden_nfa

pytestmark = pytest.mark.gpu

from admatch_amd import engine as E  # noqa: E402

import oracle  # noqa: E402  (checker only)

# (flags, selector mode) for every scan path reachable through the flags
PATHS = [
    ("plain-iNFA", E.INFA_KERNEL),
    ("TKO", E.TKO_KERNEL),
    ("AS", E.AS_KERNEL),
    ("AdMatch", E.AD_MATCHER),
    ("heuristic", E.AD_MATCHER | E.HEURISTIC_SELECTOR),
    ("learned-untrained", E.AD_MATCHER | E.LEARNED_SELECTOR),
]
ENGINE_MODES = [E.MODE_INFA, E.MODE_TKO, E.MODE_AS, E.MODE_ADMATCH, E.MODE_HEURISTIC, E.MODE_LEARNED]


def scan_all(flags, path, buf, off, variant=None):
    s = E.Scratch(flags, path).alloc_scratch()
    if variant:
        s.set_variant(*variant)
    out = E.unpack(*s.batched_scan_packed(buf, off))
    s.free_scratch()
    s.close()
    return out


def first_diff(got, exp):
    for i, (g, e) in enumerate(zip(got, exp)):
        if g != e:
            return i, g, e
    return None


@pytest.mark.parametrize("name,flags", PATHS, ids=[p[0] for p in PATHS])
@pytest.mark.parametrize("group", ["g0", "g6", "g7", "g9"])
def test_crs_groups(gold_dir, group, name, flags):
    path = golden_nfa(f"crs_{group}.nfa.gz", gold_dir)
    buf, off = golden_inputs("http_inputs.npz")
    exp = golden_json("crs_expected.json")[group]
    print(f"[case] crs {group} {name} n={off.size - 1}", flush=True)
    got = scan_all(flags, path, buf, off)
    assert got == exp, first_diff(got, exp)


@pytest.mark.parametrize("tko", [-1, 8])
@pytest.mark.parametrize("group", ["g7", "g9"])
def test_crs_tko_thresholds(gold_dir, group, tko):
    # the TKO partition decides how steps run, never results: the reference's
    # state_count/16 (-1) and a low count (8: up to 149 rows per byte, most
    # with negative offsets); the default (W/2) runs in every other test
    path = golden_nfa(f"crs_{group}.nfa.gz", gold_dir)
    buf, off = golden_inputs("http_inputs.npz")
    exp = golden_json("crs_expected.json")[group]
    for flags in (E.TKO_KERNEL, E.AD_MATCHER):
        print(f"[case] crs {group} tko threshold {tko} flags {flags}", flush=True)
        s = E.Scratch(flags, path).set_tko_threshold(tko).alloc_scratch()
        got = E.unpack(*s.batched_scan_packed(buf, off))
        s.close()
        assert got == exp, first_diff(got, exp)


@pytest.mark.parametrize("name,flags", PATHS, ids=[p[0] for p in PATHS])
def test_synthetic_small_vs_re(gold_dir, name, flags):
    path = golden_nfa("synth_small.nfa.gz", gold_dir)
    buf, off = golden_inputs("synth_small_inputs.npz")
    exp = golden_json("synth_small_expected.json")["expected"]
    print(f"[case] synth-small {name}", flush=True)
    got = scan_all(flags, path, buf, off)
    assert got == exp, first_diff(got, exp)


def test_rule1_known_answers(gold_dir):
    path = golden_nfa("crs_rule1.nfa", gold_dir)
    cases = [(b"x application/xml y", [1]), (b"text/xml", [1]), (b"text/xm", []), (b"", []),
             (b"application/xmlapplication/xml", [1]), (b"TEXT/XML", []), (b"texttext/xml\x00z", [1])]
    for name, flags in PATHS:
        s = E.Scratch(flags, path).alloc_scratch()
        got = s.batched_scan([c for c, _ in cases])
        assert got == [e for _, e in cases], name
        assert s.scan(b"a text/xml") == [1]
        s.close()


@pytest.mark.parametrize("cfg,n", [("small", 1024), ("medium", 2048), ("large", 384)])
def test_synthetic_configs_vs_oracle(workloads, cfg, n):
    w, path = workloads(cfg, n)
    buf, off = w.inputs.buf, w.inputs.offsets
    exp = oracle.OracleNFA(path).batch_scan(buf, off, nthreads=os.cpu_count() or 4)
    s = E.Scratch(E.AD_MATCHER, path).alloc_scratch()
    d_buf, d_off = E.DeviceArray(buf), E.DeviceArray(off)
    kinds = [(E.KIND_ENGINE, m) for m in ENGINE_MODES]
    if s.info()["plain_lds_bytes"] <= 160 * 1024:
        kinds.append((E.KIND_PLAIN, E.MODE_INFA))
    for kind, mode in kinds:
        print(f"[case] {cfg} n={n} kind={kind} mode={E.MODE_NAMES[mode]}", flush=True)
        s.scan_device(d_buf.ptr, d_off.ptr, off.size - 1, kind, mode)
        got = E.unpack(*s.collect(off.size - 1))
        assert got == exp, (kind, mode, first_diff(got, exp))
    s.close()


def test_edge_strings(workloads):
    w, path = workloads("small", 64)
    rng = np.random.default_rng(7)
    strs = [b"", b"a", b"\x00", b"\x00" * 300, bytes(rng.integers(0, 256, 1000, dtype=np.uint8)),
            b"z" * 5000, w.inputs.string(0) * 9]
    for i in range(40):  # lengths around the 256-byte window and 64-lane boundaries
        strs.append(bytes(rng.integers(32, 127, 250 + i * 3, dtype=np.uint8)))
    strs += [w.inputs.string(i) for i in range(64)]
    o = oracle.OracleNFA(path)
    exp = [o.scan(s) for s in strs]
    for name, flags in PATHS:
        got = scan_all(flags, path, *E.pack_strings(strs))
        assert got == exp, (name, first_diff(got, exp))


@pytest.mark.parametrize("wpb", [1, 2, 4])
@pytest.mark.parametrize("cap", [256, 512, 1024])
def test_variants_agree(workloads, wpb, cap):
    w, path = workloads("medium", 512)
    buf, off = w.inputs.buf, w.inputs.offsets
    exp = oracle.OracleNFA(path).batch_scan(buf, off, nthreads=os.cpu_count() or 4)
    for flags in (E.AD_MATCHER, E.AS_KERNEL, E.TKO_KERNEL):
        got = scan_all(flags, path, buf, off, variant=(wpb, cap, 64))
        assert got == exp, (wpb, cap, flags, first_diff(got, exp))


def test_graph_replay_and_append(workloads):
    w, path = workloads("small", 512)
    buf, off = w.inputs.buf, w.inputs.offsets
    exp = oracle.OracleNFA(path).batch_scan(buf, off, nthreads=4)
    s = E.Scratch(E.AD_MATCHER, path).alloc_scratch()
    d_buf, d_off = E.DeviceArray(buf), E.DeviceArray(off)
    for rep in range(4):
        s.scan_device(d_buf.ptr, d_off.ptr, off.size - 1, E.KIND_ENGINE, E.MODE_HEURISTIC, use_graph=True)
        got = E.unpack(*s.collect(off.size - 1))
        assert got == exp, rep
        # interleave eager work between replays
        assert s.batched_scan([w.inputs.string(1)]) == [exp[1]]
    s.close()


def test_large_full_size(workloads):
    # the large workload at its full size (65,536 strings, 92 MB): the engine
    # (hashed storage; AS and the learned table) equals the plain iNFA kernel
    # bit for bit, every planted string reports a rule, ids are sorted and
    # unique, and a sample holding the 64 longest strings equals the oracle
    from admatch_amd import selector
    w, path = workloads("large")
    buf, off = w.inputs.buf, w.inputs.offsets
    n = off.size - 1
    assert n == 65536
    s = E.Scratch(E.AD_MATCHER, path).alloc_scratch()
    d_buf, d_off = E.DeviceArray(buf), E.DeviceArray(off)
    print("[case] large full size, plain iNFA", flush=True)
    s.scan_device(d_buf.ptr, d_off.ptr, n, E.KIND_PLAIN, E.MODE_INFA)
    ref = E.unpack(*s.collect(n))
    s.set_variant(2, 256, 64, 1)
    for mode, wt in ((E.MODE_AS, None), (E.MODE_LEARNED, selector.load_weights())):
        s.set_selector(wt)
        print(f"[case] large full size, hashed mode={E.MODE_NAMES[mode]} trained={wt is not None}", flush=True)
        s.scan_device(d_buf.ptr, d_off.ptr, n, E.KIND_ENGINE, mode)
        got = E.unpack(*s.collect(n))
        assert got == ref, (mode, first_diff(got, ref))
    s.close()
    assert all(ref[i] for i in np.nonzero(w.inputs.planted)[0])
    assert all(r == sorted(set(r)) for r in ref)
    lens = np.diff(off)
    sample = sorted(set(np.argsort(lens)[-64:].tolist()) | set(range(0, n, 512)))
    sbuf, soff = E.pack_strings([w.inputs.string(i) for i in sample])
    exp = oracle.OracleNFA(path).batch_scan(sbuf, soff, nthreads=os.cpu_count() or 4)
    assert [ref[i] for i in sample] == exp


def _nfa_edges(path):
    """(initial, accepting, {src: [(dst, byte list)]}) of a per-rule file."""
    init, acc, edges = set(), set(), {}
    for line in open(path, newline="").read().replace("\r", "").split("\n"):
        if ":" not in line or line.startswith("#"):
            continue
        lhs, rhs = (x.strip() for x in line.split(":", 1))
        if "->" in lhs:
            a, b = (int(x) for x in lhs.split("->"))
            syms = []
            for atom in rhs.split():
                lo, _, hi = atom.partition("|")
                syms += list(range(int(lo), int(hi or lo) + 1))
            edges.setdefault(a, []).append((b, [c for c in syms if c < 256]))
        elif "initial" in rhs:
            init.add(int(lhs))
        elif "accepting" in rhs:
            acc.add(int(lhs))
    return init, acc, edges


def _accepted_string(path, rng, max_len=300):
    """A random member of the file's language: a path from an initial to an
    accepting state (it must end there: acceptance is tested on the final
    states only, and these files have no self-loop on their initial state)."""
    init, acc, edges = _nfa_edges(path)
    live = set(acc)  # states that can still reach an accepting state
    grow = True
    while grow:
        grow = False
        for a, outs in edges.items():
            if a not in live and any(b in live and syms for b, syms in outs):
                live.add(a)
                grow = True
    starts = sorted(init & live)
    if not starts:
        return None
    for _ in range(50):
        st, out = int(rng.choice(starts)), []
        while len(out) < max_len:
            if st in acc and (rng.random() < 0.3 or not edges.get(st)):
                return bytes(out)
            outs = [(b, syms) for b, syms in edges.get(st, []) if b in live and syms]
            if not outs:
                break
            b, syms = outs[int(rng.integers(len(outs)))]
            out.append(int(rng.choice(syms)))
            st = b
    return None


def test_snort_anchored_rules(tmp_path, gold_dir):
    # data/snort rules have no 0|255 self-loops: a rule accepts only a whole
    # string of its language (anchored both ends).  A merged sample of them,
    # scanned on every path, against the oracle; strings built from
    # accepting paths must report their own rule, and the same strings with
    # a byte added or dropped give the anchored answer the oracle computes
    from admatch_amd import merge
    from conftest import ref_fixture_files
    files = ref_fixture_files(gold_dir, "snort")
    rng = np.random.default_rng(11)
    pick = sorted(rng.choice(len(files), size=80, replace=False).tolist())
    chosen = sorted((files[i] for i in pick), key=lambda p: merge.natural_key(os.path.basename(p)))
    path = str(tmp_path / "snort80.nfa")
    with open(path, "w") as f:
        f.write(merge.merge_files(chosen))
    strs, own = [], []
    for rid, p in enumerate(chosen, start=1):
        for _ in range(3):
            s = _accepted_string(p, rng)
            if s is None:
                continue
            strs += [s, s + b"x", s[1:], b"GET " + s]
            own += [rid, 0, 0, 0]
    w = golden_inputs("http_inputs.npz")
    strs += [w[0][w[1][i]:w[1][i + 1]].tobytes() for i in range(0, 200, 4)]
    own += [0] * 50
    o = oracle.OracleNFA(path)
    exp = [o.scan(s) for s in strs]
    hits = [i for i, r in enumerate(own) if r]
    assert len(hits) > 100 and all(own[i] in exp[i] for i in hits)
    buf, off = E.pack_strings(strs)
    for name, flags in PATHS:
        print(f"[case] snort80 {name} n={len(strs)}", flush=True)
        got = scan_all(flags, path, buf, off)
        assert got == exp, (name, first_diff(got, exp))
    got = scan_all(E.AD_MATCHER, path, buf, off, variant=(2, 256, 64, 1))
    assert got == exp, ("hashed", first_diff(got, exp))
