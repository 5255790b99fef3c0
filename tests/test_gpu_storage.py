"""GPU parity of the engine's two state storages (LDS bitmaps, and hashed
active lists with HBM bitmaps) on hand-shaped rule sets that reach every
storage transition: word-boundary chains, TKO shift groups, and active sets
that swing between a few states and more than a thousand in one string
(list overflow in an AS step, dense steps, the return to the list form).
Every launch variant and selector mode against the CPU oracle, bit-exact."""
import os

import numpy as np
import pytest

from nfa_gen import burst_nfa, burst_strings, random_nfa, shift_nfa

pytestmark = pytest.mark.gpu

from admatch_amd import engine as E  # noqa: E402
from admatch_amd import selector  # noqa: E402

import oracle  # noqa: E402  (checker only)

ALL_MODES = [E.MODE_INFA, E.MODE_TKO, E.MODE_AS, E.MODE_ADMATCH, E.MODE_HEURISTIC, E.MODE_LEARNED]
HASHED_MODES = [E.MODE_AS, E.MODE_LEARNED]
VARIANTS = [(2, 512, 0), (1, 256, 1), (2, 512, 1), (4, 256, 1), (1, 1024, 1), (4, 1024, 1)]


def make_case(tmp_path, kind, seed):
    rng = np.random.default_rng(seed)
    if kind == "random":
        text = random_nfa(seed, S=200, n_edges=900)
        strs = [bytes(rng.choice(list(b"ab_ Z\x00"), size=int(rng.integers(0, 300)))) for _ in range(150)]
    elif kind == "rows":
        # AS table entries at their length boundaries (one destination, 2 and
        # 253 in the store, 254 in the long form, a persistent state plus one)
        # and an always-active state, so the two-byte rows exist
        lines = ["600"]
        lines += [f"0 -> {d} : 97" for d in range(1, 255)]
        lines += [f"0 -> {d} : 98" for d in range(1, 254)]
        lines += ["0 -> 1 : 99", "0 -> 2 : 99", "0 -> 5 : 100"]
        lines += [f"{d} -> {d + 300} : 120" for d in range(1, 255, 7)]
        lines += ["300 -> 300 : 0|255", "300 -> 301 : 120", "300 -> 302 : 121", "300 -> 303 : 121"]
        lines += ["500 -> 500 : 0|255", "500 -> 0 : 122", "500 -> 300 : 119", "0 : initial", "500 : initial"]
        lines += ["300 : accepting 1", "303 : accepting 2", "5 : accepting 3", "254 : accepting 4", "553 : accepting 5"]
        text = "\n".join(lines) + "\n"
        # random prefixes, and endings that leave accepting states active
        ends = [b"", b"wx", b"zd", b"za", b"zax", b"zbx", b"zb", b"wyy", b"zc"]
        strs = [rng.choice(list(b"abcdxyzw "), size=int(rng.integers(0, 200))).astype(np.uint8).tobytes()
                + ends[int(rng.integers(0, 9))] for _ in range(150)]
    elif kind == "shift":
        text = shift_nfa(seed)
        strs = [bytes(rng.choice(list(b"abcx"), size=int(rng.integers(0, 300)))) for _ in range(150)]
    else:
        text = burst_nfa(seed)
        strs = burst_strings(seed, 150)
    path = str(tmp_path / f"{kind}{seed}.nfa")
    with open(path, "w") as f:
        f.write(text)
    o = oracle.OracleNFA(path)
    return path, strs, [o.scan(s) for s in strs]


def first_diff(got, exp):
    for i, (g, e) in enumerate(zip(got, exp)):
        if g != e:
            return i, g, e
    return None


@pytest.mark.parametrize("kind,seed", [("random", 0), ("random", 1), ("shift", 0), ("burst", 0), ("burst", 1),
                                       ("rows", 0)])
def test_storages_vs_oracle(tmp_path, kind, seed):
    path, strs, exp = make_case(tmp_path, kind, seed)
    buf, off = E.pack_strings(strs)
    s = E.Scratch(E.AD_MATCHER, path).alloc_scratch()
    d_buf, d_off = E.DeviceArray(buf), E.DeviceArray(off)
    weights = np.load(selector.WEIGHTS_PATH) if os.path.exists(selector.WEIGHTS_PATH) else None
    for wpb, cap, store in VARIANTS:
        s.set_variant(wpb, cap, 64, store)
        for mode in (HASHED_MODES if store else ALL_MODES):
            for w in ([None, weights] if (mode == E.MODE_LEARNED and weights is not None) else [None]):
                s.set_selector(w)
                print(f"[case] {kind}{seed} wpb={wpb} cap={cap} store={store} mode={E.MODE_NAMES[mode]} "
                      f"trained={w is not None}", flush=True)
                s.scan_device(d_buf.ptr, d_off.ptr, len(strs), E.KIND_ENGINE, mode)
                got = E.unpack(*s.collect(len(strs)))
                assert got == exp, (wpb, cap, store, mode, first_diff(got, exp))
    s.close()


@pytest.mark.parametrize("tko", [-1, 0])
@pytest.mark.parametrize("kind,seed", [("shift", 1), ("burst", 0)])
def test_tko_thresholds_vs_oracle(tmp_path, kind, seed, tko):
    # every offset group a shift row (0) and the reference's partition (-1),
    # through every method that runs TKO steps, with both storages' dense paths
    path, strs, exp = make_case(tmp_path, kind, seed)
    buf, off = E.pack_strings(strs)
    s = E.Scratch(E.AD_MATCHER, path).set_tko_threshold(tko).alloc_scratch()
    d_buf, d_off = E.DeviceArray(buf), E.DeviceArray(off)
    for mode in (E.MODE_TKO, E.MODE_ADMATCH, E.MODE_HEURISTIC):
        print(f"[case] {kind}{seed} tko threshold {tko} mode={E.MODE_NAMES[mode]}", flush=True)
        s.scan_device(d_buf.ptr, d_off.ptr, len(strs), E.KIND_ENGINE, mode)
        got = E.unpack(*s.collect(len(strs)))
        assert got == exp, (tko, mode, first_diff(got, exp))
    s.close()


def test_hashed_graph_replay_and_growth(tmp_path):
    # captured launches keep their HBM bitmaps valid when a larger batch
    # grows the buffer (the graphs are dropped and captured again)
    path, strs, exp = make_case(tmp_path, "burst", 2)
    s = E.Scratch(E.AD_MATCHER, path).alloc_scratch()
    s.set_variant(2, 256, 64, 1)
    for n in (8, 150, 8, 150):
        buf, off = E.pack_strings(strs[:n])
        d_buf, d_off = E.DeviceArray(buf), E.DeviceArray(off)
        for rep in range(2):
            s.scan_device(d_buf.ptr, d_off.ptr, n, E.KIND_ENGINE, E.MODE_AS, use_graph=True)
            got = E.unpack(*s.collect(n))
            assert got == exp[:n], (n, rep, first_diff(got, exp[:n]))
    s.close()


@pytest.mark.parametrize("cfg,n", [("small", 1024), ("medium", 2048), ("large", 384)])
def test_hashed_synthetic_configs(workloads, cfg, n):
    w, path = workloads(cfg, n)
    buf, off = w.inputs.buf, w.inputs.offsets
    exp = oracle.OracleNFA(path).batch_scan(buf, off, nthreads=os.cpu_count() or 4)
    s = E.Scratch(E.AD_MATCHER, path).alloc_scratch()
    d_buf, d_off = E.DeviceArray(buf), E.DeviceArray(off)
    weights = np.load(selector.WEIGHTS_PATH) if os.path.exists(selector.WEIGHTS_PATH) else None
    for cap in (256, 512):
        s.set_variant(4, cap, 64, 1)
        for mode, wt in ((E.MODE_AS, None), (E.MODE_LEARNED, None), (E.MODE_LEARNED, weights)):
            s.set_selector(wt)
            print(f"[case] {cfg} n={n} cap={cap} hashed mode={E.MODE_NAMES[mode]} trained={wt is not None}",
                  flush=True)
            s.scan_device(d_buf.ptr, d_off.ptr, off.size - 1, E.KIND_ENGINE, mode)
            got = E.unpack(*s.collect(off.size - 1))
            assert got == exp, (cap, mode, first_diff(got, exp))
    s.close()


def test_pool_growth(tmp_path):
    # 2,000 always-accepting states per string overflow the initial
    # accepted-state pool (8 per string): the library grows it to fit and the
    # scan is repeated, through every entry point
    S = 2001
    lines = [str(S)] + [f"{i} : initial\n{i} -> {i} : 0|255\n{i} : accepting {1 + i % 50}" for i in range(1, S)]
    lines += ["0 -> 0 : 97"]
    path = str(tmp_path / "many.nfa")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    strs = [b"", b"a", b"xyz" * 50] * 22
    exp = [list(range(1, 51))] * len(strs)
    assert oracle.OracleNFA(path).batch_scan(*E.pack_strings(strs)) == exp
    for store in (0, 1):
        s = E.Scratch(E.AD_MATCHER, path).alloc_scratch()
        s.set_variant(2, 256, 64, store)
        assert s.batched_scan(strs) == exp, store
        buf, off = E.pack_strings(strs)
        d_buf, d_off = E.DeviceArray(buf), E.DeviceArray(off)
        for mode in (E.MODE_AS, E.MODE_HEURISTIC):
            s.scan_device(d_buf.ptr, d_off.ptr, len(strs), E.KIND_ENGINE, mode)
            assert E.unpack(*s.collect(len(strs))) == exp, (store, mode)
        s.close()


@pytest.mark.parametrize("kind,seed", [("random", 3), ("burst", 3)])
def test_without_pair_rows(tmp_path, monkeypatch, workloads, kind, seed):
    # rule sets whose two-byte rows exceed the entry limit run without them:
    # the first positions are listed before each AS step
    monkeypatch.setenv("ADM_MAX_PAIR_ENTRIES", "0")
    cases = [make_case(tmp_path, kind, seed)]
    w, mpath = workloads("medium", 1024)
    strs = [w.inputs.string(i) for i in range(w.inputs.n)]
    cases.append((mpath, strs, oracle.OracleNFA(mpath).batch_scan(w.inputs.buf, w.inputs.offsets, nthreads=8)))
    for path, strs, exp in cases:
        s = E.Scratch(E.AD_MATCHER, path).alloc_scratch()
        assert int(E.lib().ita_table(s._h, 15, None, 0)) == 0
        buf, off = E.pack_strings(strs)
        d_buf, d_off = E.DeviceArray(buf), E.DeviceArray(off)
        for wpb, cap, store in ((2, 512, 0), (2, 256, 1)):
            s.set_variant(wpb, cap, 64, store)
            for mode in ((E.MODE_AS, E.MODE_LEARNED, E.MODE_HEURISTIC) if store == 0 else HASHED_MODES):
                print(f"[case] no pair rows {os.path.basename(path)} store={store} mode={E.MODE_NAMES[mode]}",
                      flush=True)
                s.scan_device(d_buf.ptr, d_off.ptr, len(strs), E.KIND_ENGINE, mode)
                got = E.unpack(*s.collect(len(strs)))
                assert got == exp, (store, mode, first_diff(got, exp))
        s.close()
