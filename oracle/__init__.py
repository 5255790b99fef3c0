"""CPU oracle for the AdMatch scan path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this package, and only as the checker.  The
product (``admatch_amd``) never imports it.

``liboracle.so`` (built from ``oracle.c`` by ``make -C oracle``) restates the
reference's ``.nfa`` parser (src/transition_graph.cpp:82-294,369-411), its
iNFA step loop at thread_count == 1 (src/iNFA_kernel.cu:41-172) and its
accept extraction (src/iNFA_kernel.cu:286-307).  ``tko_partition`` restates
the TKO table build (src/transition_graph.cpp:298-357).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        L.orc_load.restype = vp
        L.orc_load.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.orc_destroy.argtypes = [vp]
        L.orc_state_count.restype = i32
        L.orc_state_count.argtypes = [vp]
        L.orc_info.argtypes = [vp, vp]
        L.orc_offsets.argtypes = [vp, vp]
        L.orc_transitions.argtypes = [vp, vp]
        L.orc_vectors.argtypes = [vp, vp, vp, vp]
        L.orc_accept_map.restype = i32
        L.orc_accept_map.argtypes = [vp, vp, vp]
        L.orc_final_states.argtypes = [vp, vp, i64, vp]
        L.orc_scan.restype = i32
        L.orc_scan.argtypes = [vp, vp, i64, vp, i32]
        L.orc_batch_scan.argtypes = [vp, vp, vp, i32, vp, vp, i32, i32]
        _LIB = L
    return _LIB


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class OracleNFA:
    """A parsed .nfa file plus the reference's single-threaded iNFA scan."""

    def __init__(self, path: str):
        err = ctypes.create_string_buffer(512)
        h = lib().orc_load(path.encode(), err, 512)
        if not h:
            raise ValueError(f"oracle: cannot load {path}: {err.value.decode()}")
        self._h = h
        info = np.zeros(10, dtype=np.int64)
        lib().orc_info(h, _ptr(info))
        (self.S, self.W32, self.n_transitions, self.n_wb, self.n_persistent, self.n_initial,
         self.n_accept_map, self.n_persistent_lines, self.n_raw_transitions, self.n_bad_lines) = map(int, info)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    # ---- graph tables -------------------------------------------------
    def offsets(self) -> np.ndarray:
        o = np.zeros(258, dtype=np.int64)
        lib().orc_offsets(self._h, _ptr(o))
        return o

    def transitions(self) -> np.ndarray:
        """(T, 2) uint32 (src, dst), per-symbol blocks sorted by (src, dst)."""
        t = np.zeros((max(self.n_transitions, 1), 2), dtype=np.uint32)
        lib().orc_transitions(self._h, _ptr(t))
        return t[: self.n_transitions]

    def vectors(self):
        """(init, persistent, accept) as uint32 word arrays."""
        a = [np.zeros(self.W32, dtype=np.uint32) for _ in range(3)]
        lib().orc_vectors(self._h, *(_ptr(x) for x in a))
        return tuple(a)

    def accept_map(self):
        n = lib().orc_accept_map(self._h, None, None)
        off = np.zeros(self.S + 1, dtype=np.int32)
        rules = np.zeros(max(n, 1), dtype=np.int32)
        lib().orc_accept_map(self._h, _ptr(off), _ptr(rules))
        return off, rules[:n]

    # ---- scanning ----------------------------------------------------
    def final_states(self, data: bytes) -> np.ndarray:
        buf = np.frombuffer(data, dtype=np.uint8) if data else np.zeros(1, np.uint8)
        out = np.zeros(self.W32, dtype=np.uint32)
        lib().orc_final_states(self._h, _ptr(buf), len(data), _ptr(out))
        return out

    def scan(self, data: bytes) -> list:
        buf = np.frombuffer(data, dtype=np.uint8) if data else np.zeros(1, np.uint8)
        cap = 256
        while True:
            out = np.zeros(cap, dtype=np.int32)
            n = lib().orc_scan(self._h, _ptr(buf), len(data), _ptr(out), cap)
            if n <= cap:
                return out[:n].tolist()
            cap = n

    def batch_scan(self, buf: np.ndarray, offsets: np.ndarray, nthreads: int = 1, per_cap: int = 64):
        """Scan strings buf[offsets[i]:offsets[i+1]] -> list of sorted rule lists."""
        buf = np.ascontiguousarray(buf, dtype=np.uint8)
        if buf.size == 0:
            buf = np.zeros(1, np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        n = offsets.size - 1
        while True:
            counts = np.zeros(max(n, 1), dtype=np.int32)
            rules = np.zeros(max(n, 1) * per_cap, dtype=np.int32)
            lib().orc_batch_scan(self._h, _ptr(buf), _ptr(offsets), n, _ptr(counts), _ptr(rules), per_cap, nthreads)
            mx = int(counts[:n].max()) if n else 0
            if mx <= per_cap:
                break
            per_cap = mx
        return [rules[i * per_cap: i * per_cap + int(counts[i])].tolist() for i in range(n)]


def tko_partition(nfa: OracleNFA, max_k: int = 1024, threshold: int = -1):
    """Restatement of the TKO table build, src/transition_graph.cpp:197-245,298-357.

    Per byte symbol c, the transitions on c plus one (s, s) per persistent
    state (the 0|255 self-loop, :197-210) are grouped by offset d = dst - src.
    Groups are ranked by size, largest first; a group is a top-k group while
    its size exceeds state_count/16 (the reference's (state_count/8)/
    sizeof(unsigned short), :308) and fewer than max_k groups are taken.  All
    other transitions, and every word-boundary transition, form the leftover
    list.  Ties in size are ranked by ascending offset (the reference's
    std::sort leaves tie order unspecified).

    threshold >= 0 replaces state_count/16 (the engine's tuned partitions;
    the grouping rule is the same).

    Returns (k_per_symbol[257], {c: [(offset, sorted source array)]},
    leftover {c: sorted unique (src, dst) array}).
    """
    off = nfa.offsets()
    tr = nfa.transitions()
    _, persis, _ = nfa.vectors()
    bits = np.unpackbits(persis.view(np.uint8), bitorder="little")[: nfa.S]
    pstates = np.nonzero(bits)[0].astype(np.int64)
    thr = nfa.S // 16 if threshold < 0 else threshold
    k_per = np.zeros(257, dtype=np.int64)
    groups, leftover = {}, {}
    for c in range(257):
        t = tr[off[c]: off[c + 1]].astype(np.int64)
        if c == 256:
            leftover[c] = t
            continue
        if pstates.size:
            # add the self-loops, then sort + unique by (src, dst) as one key
            t = np.concatenate([t, np.stack([pstates, pstates], axis=1)])
            key = np.unique(t[:, 0] * np.int64(nfa.S) + t[:, 1])
            t = np.stack([key // nfa.S, key % nfa.S], axis=1)
        d = t[:, 1] - t[:, 0]
        # groups by offset, ranked by size (largest first), ties by offset
        offs, inv, cnt = np.unique(d, return_inverse=True, return_counts=True)
        ranked = np.lexsort((offs, -cnt))
        take = ranked[:max_k][cnt[ranked[:max_k]] > thr]
        by_group = np.argsort(inv, kind="stable")  # t is sorted by source, so each group's sources stay sorted
        starts = np.concatenate([[0], np.cumsum(cnt)])
        sel = [(int(offs[g]), t[by_group[starts[g]:starts[g + 1]], 0]) for g in take]
        k_per[c] = len(sel)
        groups[c] = sel
        lo = t[~np.isin(inv, take)]
        if lo.size:
            order = np.lexsort((lo[:, 1], lo[:, 0]))
            lo = lo[order]
        leftover[c] = lo
    return k_per, groups, leftover
