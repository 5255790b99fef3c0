"""GPU parity on rule sets compiled by the regex front end: word-boundary
transitions, anchored starts and ends, case folding and empty matches at
scale, every scan path and both engine storages against the oracle (and
Python's re on short strings)."""
import numpy as np
import pytest

from admatch_amd import engine as E
from admatch_amd import regex as R
from admatch_amd import synth
from regex_gen import random_pattern

pytestmark = pytest.mark.gpu

import oracle  # noqa: E402  (checker only)
from test_regex import expected  # noqa: E402

PATHS = [("plain-iNFA", E.INFA_KERNEL), ("TKO", E.TKO_KERNEL), ("AS", E.AS_KERNEL), ("AdMatch", E.AD_MATCHER),
         ("learned", E.AD_MATCHER | E.LEARNED_SELECTOR)]


def compiled_set(tmp_path, n_rules, seed):
    rng = np.random.default_rng(seed)
    pats = []
    while len(pats) < n_rules:
        p = random_pattern(rng)
        try:
            R.compile_rules([p])
        except R.RegexError:
            continue
        pats.append(p)
    text, S = R.compile_rules(pats)
    path = str(tmp_path / f"re{seed}.nfa")
    with open(path, "w") as f:
        f.write(text)
    return pats, path


def inputs(seed, n):
    rng = np.random.default_rng(seed)
    alpha = np.frombuffer(b"abcxyzABCXYZ019._-/ \n", dtype=np.uint8)
    strs = [alpha[rng.integers(0, alpha.size, int(rng.integers(0, 400)))].tobytes() for _ in range(n // 2)]
    lengths = rng.integers(200, 1500, n - len(strs))
    http = synth.make_inputs(None, n - len(strs), lengths, 0.0, seed)
    strs += [http.string(i) for i in range(http.n)]
    return strs


@pytest.mark.parametrize("seed", [0, 1])
def test_compiled_rules_every_path(tmp_path, seed):
    pats, path = compiled_set(tmp_path, 300, seed)
    strs = inputs(seed, 600)
    o = oracle.OracleNFA(path)
    exp = [o.scan(s) for s in strs]
    # the oracle agrees with re on short strings (nested repetitions make
    # re backtrack exponentially on long ones)
    short = [x[:24] for x in strs[:60]]
    assert [o.scan(x) for x in short] == expected(pats, short)
    buf, off = E.pack_strings(strs)
    for name, flags in PATHS:
        s = E.Scratch(flags, path).alloc_scratch()
        print(f"[case] regex set {seed} {name}", flush=True)
        assert E.unpack(*s.batched_scan_packed(buf, off)) == exp, name
        s.close()
    s = E.Scratch(E.AD_MATCHER, path).alloc_scratch()
    d_buf, d_off = E.DeviceArray(buf), E.DeviceArray(off)
    for store in (0, 1):
        s.set_variant(2, 256, 64, store)
        for mode in (E.MODE_AS, E.MODE_LEARNED, E.MODE_HEURISTIC):
            print(f"[case] regex set {seed} store={store} mode={E.MODE_NAMES[mode]}", flush=True)
            s.scan_device(d_buf.ptr, d_off.ptr, len(strs), E.KIND_ENGINE, mode)
            assert E.unpack(*s.collect(len(strs))) == exp, (store, mode)
    s.close()
