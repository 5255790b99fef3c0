"""Streaming scans (ita_stream_*): strings split into chunks over several
calls, many streams interleaved; after every chunk the stream's result must
equal the oracle's result for the prefix scanned so far (word boundaries at
chunk edges included), bit-exact."""
import numpy as np
import pytest

from nfa_gen import burst_nfa, burst_strings, random_nfa

from admatch_amd import engine as E
from admatch_amd import regex as R
from regex_gen import random_pattern

pytestmark = pytest.mark.gpu

import oracle  # noqa: E402  (checker only)


def rule_set(tmp_path, kind):
    rng = np.random.default_rng(5)
    if kind == "random":
        text = random_nfa(3, S=200, n_edges=900)
        strs = [bytes(rng.choice(list(b"ab_ Z\x00"), size=int(rng.integers(0, 200)))) for _ in range(40)]
    elif kind == "burst":
        text = burst_nfa(2)
        strs = burst_strings(3, 40, max_len=300)
    else:
        pats = []
        while len(pats) < 120:
            p = random_pattern(rng)
            try:
                R.compile_rules([p])
            except R.RegexError:
                continue
            pats.append(p)
        text, _ = R.compile_rules(pats)
        alpha = np.frombuffer(b"abcxyz019._-/ \n", dtype=np.uint8)
        strs = [alpha[rng.integers(0, alpha.size, int(rng.integers(0, 200)))].tobytes() for _ in range(40)]
    path = str(tmp_path / f"{kind}.nfa")
    with open(path, "w") as f:
        f.write(text)
    return path, strs


@pytest.mark.parametrize("kind", ["random", "burst", "regex"])
@pytest.mark.parametrize("flags,store", [(E.AD_MATCHER, 0), (E.AD_MATCHER | E.LEARNED_SELECTOR, 1),
                                         (E.INFA_KERNEL, 0)])
def test_chunked_streams_match_prefixes(tmp_path, kind, flags, store):
    path, strs = rule_set(tmp_path, kind)
    o = oracle.OracleNFA(path)
    rng = np.random.default_rng(11)
    # cut every string into 1..6 chunks (some empty)
    chunks = []
    for s in strs:
        cuts = sorted(rng.integers(0, len(s) + 1, int(rng.integers(0, 6))).tolist())
        pieces = [s[a:b] for a, b in zip([0] + cuts, cuts + [len(s)])]
        chunks.append(pieces)
    sc = E.Scratch(flags, path).alloc_scratch()
    sc.set_variant(2, 256, 64, store)
    st = E.Streams(sc, len(strs))
    pos = [0] * len(strs)
    done = [0] * len(strs)
    calls = 0
    while any(pos[i] < len(chunks[i]) for i in range(len(strs))):
        # a random subset of the unfinished streams, each once
        live = [i for i in range(len(strs)) if pos[i] < len(chunks[i])]
        pick = [i for i in live if rng.random() < 0.7] or live[:1]
        got = st.scan([chunks[i][pos[i]] for i in pick], pick)
        for i, g in zip(pick, got):
            done[i] += len(chunks[i][pos[i]])
            pos[i] += 1
            assert g == o.scan(strs[i][:done[i]]), (kind, i, pos[i], calls)
        calls += 1
    # a reset stream starts from the initial states again
    st.reset(0)
    assert st.scan([strs[1]], [0]) == [o.scan(strs[1])]
    st.close()
    sc.close()


def test_stream_argument_errors(tmp_path):
    path, strs = rule_set(tmp_path, "random")
    sc = E.Scratch(E.AD_MATCHER, path).alloc_scratch()
    st = E.Streams(sc, 4)
    with pytest.raises(E.AdmatchError):
        st.scan([b"a", b"b"], [1, 1])
    with pytest.raises(E.AdmatchError):
        st.scan([b"a"], [4])
    assert st.scan([], []) == []
    st.close()
    sc.close()


def no_boundary_rule_set(tmp_path):
    # an always-active start, chains and a persistent gap, no word boundaries
    lines = ["12", "0 -> 0 : 0|255", "0 : initial", "0 -> 1 : 97", "1 -> 2 : 98", "2 -> 2 : 0|255", "2 -> 3 : 99",
             "3 -> 3 : 0|255", "3 : accepting 1", "0 -> 4 : 120", "4 -> 5 : 121", "5 -> 8 : 120", "8 -> 8 : 0|255",
             "8 : accepting 2", "0 -> 6 : 97", "6 -> 6 : 97", "6 -> 7 : 98", "7 : accepting 3"]
    path = str(tmp_path / "nowb.nfa")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    rng = np.random.default_rng(9)
    strs = [rng.choice(list(b"abcxy "), size=int(rng.integers(0, 300))).astype(np.uint8).tobytes() for _ in range(60)]
    return path, strs


@pytest.mark.parametrize("kind", ["nowb", "random", "regex"])
def test_lean_and_full_kernels_agree(tmp_path, kind):
    # a whole-string scan of a rule set without word-boundary transitions (and
    # with two-byte rows) runs the lean kernel, a stream always the full one;
    # each string scanned as one chunk must give the oracle's result either way
    path, strs = no_boundary_rule_set(tmp_path) if kind == "nowb" else rule_set(tmp_path, kind)
    o = oracle.OracleNFA(path)
    exp = [o.scan(s) for s in strs]
    sc = E.Scratch(E.AD_MATCHER, path).alloc_scratch()
    info = sc.info()
    for mode in (E.MODE_AS, E.MODE_TKO, E.MODE_ADMATCH):
        print(f"[case] {kind} mode={E.MODE_NAMES[mode]}", flush=True)
        buf, off = E.pack_strings(strs)
        d_buf, d_off = E.DeviceArray(buf), E.DeviceArray(off)
        sc.scan_device(d_buf.ptr, d_off.ptr, len(strs), E.KIND_ENGINE, mode)
        whole = E.unpack(*sc.collect(len(strs)))
        kmode, store = sc.last_kernel()
        assert (kmode & 0xff, store) == (mode, 0)
        lean = info["wb_transitions"] == 0 and sc.table(14).size > 0
        assert lean == (kind == "nowb")
        assert bool(kmode & E.LEAN_BIT) == lean, (kind, kmode)
        assert whole == exp
    st = E.Streams(sc, len(strs))
    got = st.scan(strs, list(range(len(strs))))
    kmode, _ = sc.last_kernel()
    assert not (kmode & E.LEAN_BIT)
    assert got == exp
    st.close()
    sc.close()
