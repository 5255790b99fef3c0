"""Python binding of the scan engine's C ABI (include/admatch.h).

Mirrors the reference's API (include/ITA_regex.h:30-68): ``ItaScratch`` /
``AdScratch`` are built from a flag word and a ``.nfa`` path,
``alloc_scratch`` / ``free_scratch`` move the tables to and from the GPU,
``scan`` takes one string (length ``strlen``) and ``batched_scan`` many,
returning sorted accepted rule ids per string.  Calls go straight to
``libadmatch.so``; there is no fallback: without the built library (or
without a GPU, for the calls that need one) they raise.
"""
from __future__ import annotations

import ctypes
import os
from typing import List, Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libadmatch.so")

# flag bits (include/admatch.h; 1..32 as the reference's ITA_regex.h:10-24)
INFA_KERNEL = 1
AS_KERNEL = 1 << 1
TKO_KERNEL = 1 << 2
PROFILER_MODE = 1 << 3
SHOW_RESULTS = 1 << 4
AD_MATCHER = 1 << 5
LEARNED_SELECTOR = 1 << 6
HEURISTIC_SELECTOR = 1 << 7

MODE_INFA, MODE_TKO, MODE_AS, MODE_ADMATCH, MODE_HEURISTIC, MODE_LEARNED, MODE_LOG = range(7)
KIND_PLAIN, KIND_ENGINE = 0, 1
# bits of the engine kernel's mode argument (last_kernel): lean and narrow instances
LEAN_BIT, NARROW_BIT = 0x100, 0x200
MODE_NAMES = {MODE_INFA: "infa", MODE_TKO: "tko", MODE_AS: "as", MODE_ADMATCH: "admatch",
              MODE_HEURISTIC: "heuristic", MODE_LEARNED: "learned", MODE_LOG: "log"}

ITA_OK, ITA_ERR_CAPACITY, ITA_ERR_POOL = 0, -5, -8
MLP_PARAMS = 5443
LOG_RECORD = 22

EXPORTED = [
    "ita_create", "ita_alloc", "ita_free", "ita_destroy", "ita_scan", "ita_batched_scan", "ita_scan_device",
    "ita_collect", "ita_last_stats", "ita_set_variant", "ita_get_variant", "ita_set_selector", "ita_set_log",
    "ita_fetch_log", "ita_info", "ita_table", "ita_last_error", "ita_selector_train", "ita_selector_grad",
    "ita_selector_forward", "ita_serve_trace", "ita_stream_create", "ita_stream_reset", "ita_stream_scan",
    "ita_stream_destroy", "ita_last_results", "ita_scan_device_noalloc", "ita_set_policy", "ita_serve_ring",
    "ita_last_times", "ita_set_tko_threshold", "ita_last_kernel",
]

_LIB = None


class AdmatchError(RuntimeError):
    def __init__(self, rc: int, msg: str):
        super().__init__(f"admatch error {rc}: {msg}")
        self.rc = rc


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, u16 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint16
        sig = {
            "ita_create": (i32, [u16, ctypes.c_char_p, ctypes.POINTER(vp)]),
            "ita_alloc": (i32, [vp]),
            "ita_free": (i32, [vp]),
            "ita_destroy": (i32, [vp]),
            "ita_scan": (i32, [vp, ctypes.c_char_p, vp, i32, vp]),
            "ita_batched_scan": (i32, [vp, vp, vp, i32, vp, vp, i64]),
            "ita_scan_device": (i32, [vp, vp, vp, i32, i32, i32, vp, i32]),
            "ita_collect": (i32, [vp, i32, vp, vp, i64]),
            "ita_scan_device_noalloc": (i32, [vp, vp, vp, i32, i32, i32, vp, vp, vp, vp, i64, vp]),
            "ita_last_stats": (i32, [vp, vp, vp]),
            "ita_last_times": (i32, [vp, vp]),
            "ita_last_kernel": (i32, [vp, vp]),
            "ita_set_variant": (i32, [vp, i32, i32, i32, i32]),
            "ita_get_variant": (i32, [vp, vp, vp, vp, vp]),
            "ita_set_tko_threshold": (i32, [vp, i32]),
            "ita_set_selector": (i32, [vp, vp, i32, i32]),
            "ita_set_policy": (i32, [vp, vp, i32]),
            "ita_stream_create": (i32, [vp, i32, ctypes.POINTER(vp)]),
            "ita_stream_reset": (i32, [vp, i32]),
            "ita_stream_scan": (i32, [vp, vp, vp, vp, i32, vp, vp, i64]),
            "ita_stream_destroy": (i32, [vp]),
            "ita_last_results": (i64, [vp, vp, i64]),
            "ita_set_log": (i32, [vp, i64, i32]),
            "ita_fetch_log": (i64, [vp, vp, i64]),
            "ita_info": (i32, [vp, vp]),
            "ita_table": (i64, [vp, i32, vp, i64]),
            "ita_last_error": (ctypes.c_char_p, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _err(rc: int):
    if rc != ITA_OK:
        raise AdmatchError(rc, lib().ita_last_error().decode(errors="replace"))


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def pack_strings(strings: Sequence[bytes]) -> Tuple[np.ndarray, np.ndarray]:
    lens = np.fromiter((len(s) for s in strings), dtype=np.int64, count=len(strings))
    offsets = np.zeros(len(strings) + 1, dtype=np.int64)
    np.cumsum(lens, out=offsets[1:])
    buf = np.frombuffer(b"".join(strings), dtype=np.uint8) if strings else np.zeros(0, np.uint8)
    return buf, offsets


class Scratch:
    """A compiled rule set plus its device residency (reference ita_scratch)."""

    def __init__(self, flags: int, nfa_path: str):
        h = ctypes.c_void_p()
        _err(lib().ita_create(flags, os.fsencode(nfa_path), ctypes.byref(h)))
        self._h = h
        self.flags = flags

    # -- lifecycle ------------------------------------------------------
    def alloc_scratch(self):
        _err(lib().ita_alloc(self._h))
        return self

    def free_scratch(self):
        _err(lib().ita_free(self._h))

    def close(self):
        if getattr(self, "_h", None):
            lib().ita_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    # -- scanning -------------------------------------------------------
    def scan(self, text: bytes) -> List[int]:
        cap = 1024
        while True:
            out = np.zeros(cap, dtype=np.int32)
            n = ctypes.c_int32()
            rc = lib().ita_scan(self._h, text, _p(out), cap, ctypes.byref(n))
            if rc == ITA_ERR_CAPACITY and n.value > cap:  # output too small: the full count is in n
                cap = n.value
                continue
            _err(rc)
            return out[: n.value].tolist()

    def batched_scan_packed(self, buf: np.ndarray, offsets: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
        """Scan buf[offsets[i]:offsets[i+1]]; returns (out_offsets[n+1], ids)."""
        buf = np.ascontiguousarray(buf, dtype=np.uint8)
        offsets = np.asarray(offsets, dtype=np.int64)
        n = offsets.size - 1
        base = buf.ctypes.data
        ptrs = (base + offsets[:-1]).astype(np.uint64)
        lens = np.diff(offsets).astype(np.int32)
        out_off = np.zeros(n + 1, dtype=np.int32)
        cap = max(4 * n, 1024)
        while True:
            ids = np.zeros(cap, dtype=np.int32)
            rc = lib().ita_batched_scan(self._h, _p(ptrs), _p(lens), n, _p(out_off), _p(ids), cap)
            if rc == ITA_ERR_CAPACITY and int(out_off[-1]) > cap:  # out_ids too small: offsets hold the total
                cap = int(out_off[-1])
                continue
            _err(rc)
            return out_off, ids[: out_off[-1]]

    def batched_scan(self, strings: Sequence[bytes]) -> List[List[int]]:
        buf, offsets = pack_strings(strings)
        return unpack(*self.batched_scan_packed(buf, offsets))

    # -- device-resident path (bench, serving) ---------------------------
    def scan_device(self, d_text: int, d_offsets: int, n: int, kind: int = KIND_ENGINE, mode: int = MODE_ADMATCH,
                    stream: Optional[int] = None, use_graph: bool = False):
        self._last_scan = (d_text, d_offsets, n, kind, mode, stream, int(use_graph))
        _err(lib().ita_scan_device(self._h, *self._last_scan))

    def scan_device_noalloc(self, d_text: int, d_offsets: int, n: int, kind: int, mode: int, d_counts: int,
                            d_starts: int, d_states: int, states_cap: int, d_status: int,
                            stream: Optional[int] = None):
        """ita_scan_device_noalloc: enqueues the scan, results in the given
        device buffers (accepted states per string; see include/admatch.h)."""
        _err(lib().ita_scan_device_noalloc(self._h, d_text, d_offsets, n, kind, mode, stream, d_counts, d_starts,
                                           d_states, states_cap, d_status))

    def accept_map(self) -> Tuple[np.ndarray, np.ndarray]:
        """(offsets[S+1], rule ids): the rules of state s are ids[off[s]:off[s+1]]."""
        return self.table(21), self.table(22)

    def collect(self, n: int) -> Tuple[np.ndarray, np.ndarray]:
        """Results of the last scan_device; a pool overflow (the library grows
        the pool to fit) repeats that scan once."""
        out_off = np.zeros(n + 1, dtype=np.int32)
        cap = max(4 * n, 1024)
        rescanned = False
        while True:
            ids = np.zeros(cap, dtype=np.int32)
            rc = lib().ita_collect(self._h, n, _p(out_off), _p(ids), cap)
            if rc == ITA_ERR_CAPACITY and int(out_off[-1]) > cap:
                cap = int(out_off[-1])
                continue
            if rc == ITA_ERR_POOL and not rescanned and getattr(self, "_last_scan", None):
                rescanned = True
                _err(lib().ita_scan_device(self._h, *self._last_scan))
                continue
            _err(rc)
            return out_off, ids[: out_off[-1]]

    def last_stats(self):
        ms = ctypes.c_float()
        steps = np.zeros(3, dtype=np.uint64)
        _err(lib().ita_last_stats(self._h, ctypes.byref(ms), _p(steps)))
        return ms.value, steps.tolist()

    def last_kernel(self):
        """(mode argument, storage) of the last scan's kernel: the mode plus
        LEAN_BIT / NARROW_BIT, -1 for the plain kernel; storage 1 = hashed."""
        t = np.zeros(2, dtype=np.int32)
        _err(lib().ita_last_kernel(self._h, _p(t)))
        return int(t[0]), int(t[1])

    def last_times(self) -> dict:
        """ms of the last scan: launch sequence, engine kernel alone, and the
        host-strings path's phases."""
        t = np.zeros(6, dtype=np.float32)
        _err(lib().ita_last_times(self._h, _p(t)))
        return dict(zip(("sequence", "engine_kernel", "alloc", "h2d", "d2h", "map"), map(float, t)))

    # -- configuration ---------------------------------------------------
    def set_variant(self, wpb: int = 2, cap: int = 512, plain_bs: int = 64, store: int = 0):
        """store 0: LDS bitmaps; 1: hashed active lists (AS and learned modes)."""
        _err(lib().ita_set_variant(self._h, wpb, cap, plain_bs, store))

    def set_tko_threshold(self, threshold: int):
        """TKO shift-group threshold: -2 auto (W/2, the default), -1 the
        reference's state_count/16, >= 0 that count.  Before alloc_scratch."""
        _err(lib().ita_set_tko_threshold(self._h, threshold))
        return self

    def variant(self):
        """(wpb, cap, plain_bs, store)"""
        v = [ctypes.c_int32() for _ in range(4)]
        _err(lib().ita_get_variant(self._h, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def set_selector(self, weights: Optional[np.ndarray], every: int = 0):
        if weights is None:
            _err(lib().ita_set_selector(self._h, None, 0, every))
            return
        w = np.ascontiguousarray(weights, dtype=np.float32)
        _err(lib().ita_set_selector(self._h, _p(w), w.size, every))

    def set_policy(self, table: np.ndarray):
        """Installs a policy table (512 u32) derived from the selector's network."""
        t = np.ascontiguousarray(table, dtype=np.uint32)
        _err(lib().ita_set_policy(self._h, _p(t), t.size))

    def policy_table(self) -> np.ndarray:
        """The learned selector's table (512 u32) of the current weights."""
        return self.table(16)

    def set_log(self, cap_records: int, stride: int):
        _err(lib().ita_set_log(self._h, cap_records, stride))

    def fetch_log(self, cap_records: int) -> np.ndarray:
        out = np.zeros((cap_records, LOG_RECORD), dtype=np.float32)
        n = lib().ita_fetch_log(self._h, _p(out), cap_records)
        if n < 0:
            _err(int(n))
        return out[:n]

    def info(self) -> dict:
        v = np.zeros(11, dtype=np.int64)
        _err(lib().ita_info(self._h, _p(v)))
        keys = ["states", "words", "transitions", "wb_transitions", "tko_groups", "tko_threshold", "as_entries",
                "engine_lds_bytes", "plain_lds_bytes", "cus", "hash"]
        return dict(zip(keys, map(int, v)))

    def table(self, which: int) -> np.ndarray:
        dt = {4: np.uint64, 9: np.uint64, 1: np.uint32, 6: np.uint32, 7: np.uint32, 8: np.uint32, 10: np.uint32,
              11: np.uint32, 13: np.uint32, 14: np.uint32, 15: np.uint32, 16: np.uint32, 17: np.uint32, 18: np.uint32,
              20: np.uint32}.get(which, np.int32)
        n = lib().ita_table(self._h, which, None, 0)
        if n < 0:
            _err(int(n))
        out = np.zeros(max(int(n), 1), dtype=dt)
        lib().ita_table(self._h, which, _p(out), int(n))
        return out[: int(n)]


class ItaScratch(Scratch):
    """Reference ``ita_scratch``: one kernel bit (INFA / AS / TKO)."""


class AdScratch(Scratch):
    """Reference ``ad_scratch``: the adaptive engine (AD_MATCHER implied)."""

    def __init__(self, flags: int, nfa_path: str):
        super().__init__(flags | AD_MATCHER, nfa_path)


class Streams:
    """Strings scanned in chunks over several calls (ita_stream_*): chunk i of
    a call continues stream ids[i]; each result is that stream's accepted
    rule ids as if it ended after the chunk."""

    def __init__(self, scratch: "Scratch", n_streams: int):
        self._scratch = scratch  # keeps the handle alive
        self._s = ctypes.c_void_p()
        _err(lib().ita_stream_create(scratch._h, n_streams, ctypes.byref(self._s)))

    def reset(self, stream: int):
        _err(lib().ita_stream_reset(self._s, stream))

    def scan(self, chunks: Sequence[bytes], ids: Sequence[int]) -> List[List[int]]:
        buf, offsets = pack_strings(chunks)
        n = len(chunks)
        ptrs = (buf.ctypes.data + offsets[:-1]).astype(np.uint64)
        lens = np.diff(offsets).astype(np.int32)
        sid = np.ascontiguousarray(ids, dtype=np.int32)
        out_off = np.zeros(n + 1, dtype=np.int32)
        cap = max(4 * n, 1024)
        ids_out = np.zeros(cap, dtype=np.int32)
        rc = lib().ita_stream_scan(self._s, _p(ptrs), _p(lens), _p(sid), n, _p(out_off), _p(ids_out), cap)
        if rc == ITA_ERR_CAPACITY:  # the streams advanced: fetch the kept results
            ids_out = np.zeros(int(out_off[-1]), dtype=np.int32)
            lib().ita_last_results(self._scratch._h, _p(ids_out), ids_out.size)
        else:
            _err(rc)
        return unpack(out_off, ids_out[: out_off[-1]])

    def close(self):
        if self._s:
            lib().ita_stream_destroy(self._s)
            self._s = ctypes.c_void_p()

    def __del__(self):
        self.close()


def unpack(out_off: np.ndarray, ids: np.ndarray) -> List[List[int]]:
    ids = ids.tolist()
    o = out_off.tolist()
    return [ids[o[i]: o[i + 1]] for i in range(len(o) - 1)]


# ---------------------------------------------------------------- device memory
class Hip:
    """Minimal HIP runtime access for device buffers (no torch needed)."""

    _rt = None

    @classmethod
    def rt(cls):
        if cls._rt is None:
            for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
                try:
                    cls._rt = ctypes.CDLL(name)
                    break
                except OSError:
                    continue
            if cls._rt is None:
                raise RuntimeError("libamdhip64.so not found")
            r = cls._rt
            r.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
            r.hipFree.argtypes = [ctypes.c_void_p]
            r.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            r.hipDeviceSynchronize.argtypes = []
            r.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
            r.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
            r.hipStreamDestroy.argtypes = [ctypes.c_void_p]
            r.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
            r.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
            r.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
            r.hipEventSynchronize.argtypes = [ctypes.c_void_p]
            r.hipEventQuery.argtypes = [ctypes.c_void_p]
            r.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
            r.hipEventDestroy.argtypes = [ctypes.c_void_p]
        return cls._rt

    @classmethod
    def check(cls, rc: int, what: str):
        if rc != 0:
            raise RuntimeError(f"{what} failed with hipError {rc}")

    @classmethod
    def device_count(cls) -> int:
        n = ctypes.c_int(0)
        rc = cls.rt().hipGetDeviceCount(ctypes.byref(n))
        return n.value if rc == 0 else 0


class DeviceArray:
    """A device allocation filled from a numpy array (hipMalloc + hipMemcpy)."""

    def __init__(self, host: np.ndarray):
        host = np.ascontiguousarray(host)
        self.nbytes = max(int(host.nbytes), 1)
        p = ctypes.c_void_p()
        Hip.check(Hip.rt().hipMalloc(ctypes.byref(p), self.nbytes), "hipMalloc")
        self.ptr = p.value
        if host.nbytes:
            Hip.check(Hip.rt().hipMemcpy(self.ptr, host.ctypes.data, host.nbytes, 1), "hipMemcpy H2D")

    def to_host(self, out: np.ndarray) -> np.ndarray:
        """Copies the first out.nbytes bytes back into out (hipMemcpy D2H)."""
        if out.nbytes:
            Hip.check(Hip.rt().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def free(self):
        if getattr(self, "ptr", None):
            Hip.rt().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()
