"""Small hand-shaped NFAs for tests (written as .nfa text).

* random_nfa: arbitrary transitions, ranges, persistent states, word-boundary
  transitions (including chains in both id orders), several accept rules.
* shift_nfa: many transitions per (symbol, offset) for the offsets that
  stress the TKO word shifts (0, +-1, +-31, +-32, +-33, +-63, +-64, +-65), so
  each becomes a top-k group (size > S/16).
* burst_nfa + burst_strings: an always-active start state whose rows reach
  S/26 states per letter, most of which stay active while letters continue
  and die at other bytes, so the active set swings between a few states and
  more than a thousand within one string (list overflow, dense steps and the
  return to the list form), plus word-boundary transitions.
"""
from __future__ import annotations

import numpy as np

SHIFT_OFFSETS = [0, 1, -1, 31, -31, 32, -32, 33, -33, 63, -63, 64, -64, 65, -65, 127, -129]


def _atom(lo, hi):
    return str(lo) if lo == hi else f"{lo}|{hi}"


def random_nfa(seed: int, S: int = 40, n_edges: int = 120, alphabet: bytes = b"ab_ ",
               wb_frac: float = 0.15, persist_frac: float = 0.1, n_rules: int = 5) -> str:
    rng = np.random.default_rng(seed)
    lines = ["# random", str(S)]
    for s in rng.choice(S, size=max(1, S // 10), replace=False):
        lines.append(f"{s} : initial")
    for _ in range(n_edges):
        a, b = (int(x) for x in rng.integers(0, S, 2))
        r = rng.random()
        if r < wb_frac:
            lines.append(f"{a} -> {b} : 256")
        elif r < wb_frac + 0.1:
            lo = int(rng.integers(0, 200))
            lines.append(f"{a} -> {b} : {_atom(lo, lo + int(rng.integers(0, 56)))}")
        else:
            syms = rng.choice(list(alphabet), size=int(rng.integers(1, 3)), replace=False)
            lines.append(f"{a} -> {b} : " + " ".join(str(int(c)) for c in syms))
    for s in rng.choice(S, size=max(1, int(S * persist_frac)), replace=False):
        lines.append(f"{s} -> {s} : 0|255")
    # word-boundary chains, both id orders (the pass runs in (src,dst) order)
    for _ in range(3):
        a, b, c = (int(x) for x in rng.choice(S, 3, replace=False))
        lines += [f"{a} -> {b} : 256", f"{b} -> {c} : 256"]
    accepts = rng.choice(S, size=min(S, 2 * n_rules), replace=False)
    for k, s in enumerate(accepts):
        rules = " ".join(str(1 + (k + j) % n_rules) for j in range(1 + k % 2))
        lines.append(f"{s} : accepting {rules}")
    return "\n".join(lines) + "\n"


def shift_nfa(seed: int, S: int = 600, alphabet: bytes = b"abc") -> str:
    rng = np.random.default_rng(seed)
    lines = ["# shifts", str(S)]
    for s in rng.choice(S, size=S // 3, replace=False):
        lines.append(f"{s} : initial")
    thr = S // 16
    for d in SHIFT_OFFSETS:
        c = int(alphabet[int(rng.integers(len(alphabet)))])
        lo, hi = max(0, -d), min(S, S - d)
        srcs = rng.choice(np.arange(lo, hi), size=min(hi - lo, thr + 5 + int(rng.integers(0, 10))), replace=False)
        for s in srcs:
            lines.append(f"{int(s)} -> {int(s) + d} : {c}")
    for _ in range(200):  # leftover noise
        a, b = (int(x) for x in rng.integers(0, S, 2))
        lines.append(f"{a} -> {b} : {int(alphabet[int(rng.integers(len(alphabet)))])}")
    for s in rng.choice(S, size=S // 20, replace=False):
        lines.append(f"{s} -> {s} : 0|255")
    for k, s in enumerate(rng.choice(S, size=30, replace=False)):
        lines.append(f"{s} : accepting {k + 1}")
    return "\n".join(lines) + "\n"


def burst_nfa(seed: int, S: int = 3000, n_rules: int = 50) -> str:
    rng = np.random.default_rng(seed)
    lines = ["# burst", str(S), "0 : initial", "0 -> 0 : 0|255"]
    for i in range(1, S):
        lines.append(f"0 -> {i} : {97 + i % 26}")
        if rng.random() < 0.7:
            lines.append(f"{i} -> {i} : 97|122")
    for _ in range(S // 2):  # letter chains between the burst states
        a, b = (int(x) for x in rng.integers(1, S, 2))
        lines.append(f"{a} -> {b} : {int(rng.integers(97, 123))}")
    for _ in range(S // 50):  # digit and space transitions
        a, b = (int(x) for x in rng.integers(1, S, 2))
        lines.append(f"{a} -> {b} : {int(rng.choice([32, 48, 49, 57]))}")
    for s_ in rng.choice(np.arange(1, S), size=S // 100, replace=False):
        lines.append(f"{int(s_)} -> {int(s_)} : 0|255")
    for _ in range(20):
        a, b = (int(x) for x in rng.integers(0, S, 2))
        lines.append(f"{a} -> {b} : 256")
    for k, s_ in enumerate(rng.choice(np.arange(1, S), size=2 * n_rules, replace=False)):
        lines.append(f"{int(s_)} : accepting {1 + k % n_rules}")
    return "\n".join(lines) + "\n"


def burst_strings(seed: int, n: int, max_len: int = 400):
    """Mostly letters with occasional spaces, digits and punctuation."""
    rng = np.random.default_rng(seed)
    other = np.frombuffer(b" 019.,-_", dtype=np.uint8)
    out = []
    for _ in range(n):
        L = int(rng.integers(0, max_len + 1))
        letters = rng.integers(97, 123, L).astype(np.uint8)
        mask = rng.random(L) < 0.08
        letters[mask] = rng.choice(other, int(mask.sum()))
        out.append(letters.tobytes())
    return out


def brute_force(nfa_text: str, data: bytes):
    """Set-based simulation of the step loop, written independently of the
    oracle: dict of per-symbol edges, Python sets, word-boundary pass in
    (src, dst) order with in-place updates."""
    lines = [l.rstrip("\r") for l in nfa_text.split("\n")]
    body = [l for l in lines if l and not l.startswith("#")]
    S = int(body[0])
    init, persis, edges, wb, acc = set(), set(), {}, [], {}
    for l in body[1:]:
        lhs, rhs = [x.strip() for x in l.split(":")]
        if "->" in lhs:
            a, b = (int(x) for x in lhs.split("->"))
            for atom in rhs.split():
                lo, _, hi = atom.partition("|")
                lo = int(lo)
                hi = int(hi) if hi else lo
                if a == b and lo == 0 and hi == 255:
                    persis.add(a)
                    continue
                for c in range(lo, hi + 1):
                    if c == 256:
                        wb.append((a, b))
                    else:
                        edges.setdefault(c, set()).add((a, b))
        elif "accepting" in rhs:
            toks = rhs.split()
            if len(toks) > 1 and int(lhs) not in acc:
                acc[int(lhs)] = [int(t) for t in toks[1:]]
        elif "initial" in rhs:
            init.add(int(lhs))
    wb = sorted(set(wb))

    def word(c):
        return chr(c).isascii() and (chr(c).isalnum() or c == ord("_"))

    def wbpass(cur):
        for a, b in wb:
            if a in cur:
                cur.add(b)

    cur = set(init)
    if wb and data and word(data[0]):
        wbpass(cur)
    for i, c in enumerate(data):
        nxt = {s for s in cur if s in persis}
        for a, b in edges.get(c, ()):
            if a in cur:
                nxt.add(b)
        cur = nxt
        if wb and ((i < len(data) - 1 and word(c) != word(data[i + 1])) or (i == len(data) - 1 and word(c))):
            wbpass(cur)
    out = set()
    for s in cur:
        out.update(acc.get(s, []))
    return sorted(out)
