"""Learned per-step kernel selector: logging, GPU training, packaging.

The engine's MODE_LOG runs all three step methods (iNFA, TKO, AS) on the
same state for one sampled step in ``stride`` and records the step's 16
features and each method's cycle count.  A 16 -> 64 -> 64 -> 3 MLP is then
trained on the GPU (ita_selector_train: the project's own forward/backward
and Adam kernels) to predict log2 cycles per method; at scan time the
network is evaluated once per (active-count bucket, list validity, byte)
into a 2-bit policy table (the engine's policy_kernel) that each step reads
with one scalar load; the method with the lowest predicted cost runs.
Results do not depend on the choice: every method computes the same next
state set.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import engine as E

F, H, O = 16, 64, 3
N_PARAMS = E.MLP_PARAMS
WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "selector_weights.npy")

# The step features (device: feature() in engine_kernels.hpp): active-state
# count, the byte's transition statistics, out-degree bins of the listed
# states on the byte, the byte class, position and length, and the step
# before (its method and its cycles).
FEATURE_NAMES = [
    "log2(1+active)", "active %", "log2(1+|T_c|)", "k_c", "log2(1+leftover_c)", "empty-row share (-1: bitmap)",
    "word(c)", "c/32", "pos/len", "log2(1+row entries)", "list valid", "last=iNFA", "last=TKO", "last=AS",
    "log2(1+remaining)", "log2(1+last step cycles)",
]
POLICY_STEP_CYCLES = 1024  # the step before at a synthetic table context (kPolicyStepCycles)


def _sig():
    L = E.lib()
    vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32
    L.ita_selector_train.restype = i32
    L.ita_selector_train.argtypes = [vp, vp, vp, i64, i32, i32, ctypes.c_float, ctypes.c_uint64, vp, vp]
    L.ita_selector_grad.restype = i32
    L.ita_selector_grad.argtypes = [vp, vp, vp, i64, vp, vp, vp]
    L.ita_selector_forward.restype = i32
    L.ita_selector_forward.argtypes = [vp, i64, vp, vp]
    return L


def _p(a):
    return a.ctypes.data


@dataclass
class Dataset:
    x: np.ndarray     # [n,16] float32 features
    y: np.ndarray     # [n,3] log2 cycles (0 where not timed)
    mask: np.ndarray  # [n,3] 1 where timed
    c: Optional[np.ndarray] = None       # [n] byte of the step
    n_expl: Optional[np.ndarray] = None  # [n] explicit (listed) active count

    @property
    def n(self):
        return self.x.shape[0]

    def best(self) -> np.ndarray:
        t = np.where(self.mask > 0, self.y, np.inf)
        return t.argmin(1)


def from_log(records: np.ndarray) -> Dataset:
    """Log records (22 floats: 16 features, 3 cycle counts or -1, chosen,
    byte, explicit active count)."""
    x = records[:, :F].astype(np.float32)
    t = records[:, F:F + O].astype(np.float64)
    mask = (t > 0).astype(np.float32)
    y = np.where(t > 0, np.log2(np.maximum(t, 1.0)), 0.0).astype(np.float32)
    c = records[:, F + 4].astype(np.int64) if records.shape[1] > F + 4 else None
    ne = records[:, F + 5].astype(np.int64) if records.shape[1] > F + 5 else None
    return Dataset(x=x, y=y, mask=mask, c=c, n_expl=ne)


def collect(scratch: E.Scratch, d_buf, d_off, n: int, cap: int = 1 << 20, stride: int = 16) -> np.ndarray:
    scratch.set_log(cap, stride)
    scratch.scan_device(d_buf.ptr, d_off.ptr, n, E.KIND_ENGINE, E.MODE_LOG)
    scratch.collect(n)
    rec = scratch.fetch_log(cap)
    scratch.set_log(0, stride)
    return rec


def init_weights(seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    w = np.zeros(N_PARAMS, dtype=np.float32)
    o = 0
    for fan_in, fan_out in ((F, H), (H, H)):
        w[o:o + fan_in * fan_out] = rng.normal(0, np.sqrt(2.0 / fan_in), fan_in * fan_out)
        o += fan_in * fan_out + fan_out
    w[o:o + O * H] = rng.normal(0, np.sqrt(1.0 / H), O * H)
    return w


def fold_normalisation(w: np.ndarray, mu: np.ndarray, sd: np.ndarray) -> np.ndarray:
    """Weights trained on (x-mu)/sd -> weights that take raw x."""
    w = w.astype(np.float64).copy()
    W1 = w[:F * H].reshape(F, H)
    b1 = w[F * H:F * H + H]
    b1 -= (mu / sd) @ W1
    W1 /= sd[:, None]
    w[:F * H] = W1.ravel()
    w[F * H:F * H + H] = b1
    return w.astype(np.float32)


def train(ds: Dataset, epochs: int = 40, batch: int = 512, lr: float = 2e-3, seed: int = 0):
    """Trains on the GPU; returns (packed weights for raw features, loss history)."""
    mu = ds.x.mean(0).astype(np.float64)
    sd = ds.x.std(0).astype(np.float64)
    sd[sd < 1e-6] = 1.0
    xn = ((ds.x - mu) / sd).astype(np.float32)
    # targets centred per method for a better-conditioned start
    w = init_weights(seed)
    w[-O:] = np.array([ds.y[ds.mask[:, m] > 0, m].mean() if (ds.mask[:, m] > 0).any() else 0.0
                       for m in range(O)], dtype=np.float32)
    hist = np.zeros(epochs, dtype=np.float32)
    rc = _sig().ita_selector_train(_p(np.ascontiguousarray(xn)), _p(np.ascontiguousarray(ds.y)),
                                   _p(np.ascontiguousarray(ds.mask)), ds.n, epochs, batch, lr, seed, _p(w), _p(hist))
    E._err(rc)
    return fold_normalisation(w, mu, sd), hist


def gpu_grad(x, y, mask, w):
    g = np.zeros(N_PARAMS, dtype=np.float32)
    loss = np.zeros(1, dtype=np.float32)
    x, y, mask, w = (np.ascontiguousarray(a, dtype=np.float32) for a in (x, y, mask, w))
    E._err(_sig().ita_selector_grad(_p(x), _p(y), _p(mask), x.shape[0], _p(w), _p(g), _p(loss)))
    return float(loss[0]), g


def gpu_forward(x, w):
    x = np.ascontiguousarray(x, dtype=np.float32)
    w = np.ascontiguousarray(w, dtype=np.float32)
    out = np.zeros((x.shape[0], O), dtype=np.float32)
    E._err(_sig().ita_selector_forward(_p(x), x.shape[0], _p(w), _p(out)))
    return out


def choose(pred: np.ndarray, allow_as: np.ndarray) -> np.ndarray:
    p = pred.copy()
    p[~allow_as, 2] = np.inf
    return p.argmin(1)


def regret(ds: Dataset, choice: np.ndarray) -> dict:
    """Cycles of the chosen methods vs the per-step best, over timed steps."""
    t = np.where(ds.mask > 0, np.exp2(ds.y.astype(np.float64)), np.inf)
    picked = t[np.arange(ds.n), choice]
    best = t.min(1)
    ok = np.isfinite(picked)
    return {"picked_cycles": float(picked[ok].sum()), "best_cycles": float(best[ok].sum()),
            "accuracy": float((choice == t.argmin(1)).mean())}


def heuristic_choice(ds: Dataset, W: int) -> np.ndarray:
    """The engine's fallback rule on logged features (AdMatch_kernel.cu:87
    plus an iNFA cost term), for offline comparison."""
    active = np.rint(np.exp2(ds.x[:, 0]) - 1)
    T = np.rint(np.exp2(ds.x[:, 2]) - 1)
    K = ds.x[:, 3]
    Lft = np.rint(np.exp2(ds.x[:, 4]) - 1)
    sparse = ds.x[:, 10] > 0
    out = np.where(T <= K * W + Lft, 0, 1)
    out[sparse & (K > active)] = 2
    return out


# ---------------------------------------------------------------- policy table
# Mirrors policy_kernel (admatch_amd/csrc/kernels.hip): entry (bucket b, list
# valid v, byte c) holds the method the network picks at a representative
# step context; bits 2*(c%16) of word (2*b + v)*16 + c//16.
POLICY_BUCKETS = 16
POLICY_LEN = 1024


def policy_bucket(n_expl) -> np.ndarray:
    n = np.asarray(n_expl, dtype=np.int64)
    return np.minimum(POLICY_BUCKETS - 1, np.floor(np.log2(n + 1)).astype(np.int64))


def policy_inputs(scratch: E.Scratch):
    """Features [16*2*256, 16] (float64) of every table entry, and whether
    the entry may pick AS (list valid)."""
    info = scratch.info()
    W, S = info["words"], info["states"]
    soff = scratch.table(0).astype(np.int64)
    koff = scratch.table(2).astype(np.int64)
    loff = scratch.table(5).astype(np.int64)
    alw_total = int(scratch.table(12)[256])
    T, K, Lft = np.diff(soff)[:256], np.diff(koff)[:256], np.diff(loff)[:256]
    rows, allow = [], []
    for b in range(POLICY_BUCKETS):
        n = (0 if b == 0 else (3 << (b - 1)) - 1) + alw_total // 256
        for v in (0, 1):
            for c in range(256):
                word = chr(c).isascii() and (chr(c).isalnum() or c == ord("_"))
                rows.append([np.log2(1 + n), 100.0 * n / S, np.log2(1 + T[c]), K[c], np.log2(1 + Lft[c]),
                             0.5 if v else -1.0, float(word), float(c >> 5), 0.5, np.log2(1 + n), float(v),
                             float(v == 0), 0.0, float(v == 1), np.log2(1 + POLICY_LEN // 2),
                             np.log2(1 + POLICY_STEP_CYCLES)])
                allow.append(v == 1)
    return np.array(rows, dtype=np.float64), np.array(allow)


def pack_policy(choice: np.ndarray) -> np.ndarray:
    """Methods in entry order -> the u32[512] table."""
    t = np.zeros(POLICY_BUCKETS * 2 * 16, dtype=np.uint32)
    e = np.arange(choice.size)
    np.bitwise_or.at(t, e >> 4, (choice.astype(np.uint32) & 3) << ((e & 15) * 2).astype(np.uint32))
    return t


def policy_lookup(table: np.ndarray, n_expl, sparse, c) -> np.ndarray:
    e = (policy_bucket(n_expl) * 2 + np.asarray(sparse, dtype=np.int64)) * 256 + np.asarray(c, dtype=np.int64)
    return ((table[e >> 4] >> ((e & 15) * 2).astype(np.uint32)) & 3).astype(np.int64)


def policy_entries(ds: Dataset) -> np.ndarray:
    """Table entry of each logged step."""
    return (policy_bucket(ds.n_expl) * 2 + (ds.x[:, 10] > 0).astype(np.int64)) * 256 + ds.c.astype(np.int64)


def policy_from_log(scratch: E.Scratch, w: np.ndarray, ds: Dataset) -> np.ndarray:
    """The policy table of least expected cost under the network: for every
    entry (count bucket, list validity, byte) that logged steps fall into,
    the method whose predicted cycles, summed over those steps, are lowest
    (AS only where the list is valid); the other entries keep the network's
    choice at the synthetic step context (what ita_set_selector built).  The
    table is what the network predicts for the steps the engine meets, where
    the synthetic context fixes features (position, the method before) that
    vary within an entry."""
    scratch.set_selector(w)
    base = scratch.policy_table()
    e_all = np.arange(POLICY_BUCKETS * 2 * 256)
    choice = ((base[e_all >> 4] >> ((e_all & 15) * 2).astype(np.uint32)) & 3).astype(np.int64)
    cost = np.exp2(gpu_forward(ds.x, w).astype(np.float64))
    ent = policy_entries(ds)
    sums = np.zeros((e_all.size, O))
    np.add.at(sums, ent, cost)
    seen = np.bincount(ent, minlength=e_all.size) > 0
    valid = ((e_all >> 8) & 1) == 1
    sums[~valid, 2] = np.inf
    choice[seen] = sums[seen].argmin(1)
    return pack_policy(choice)


# a table derived for one rule set, stored with that rule set's content hash
POLICY_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "selector_policy.npz")


def load_weights(path: Optional[str] = None) -> Optional[np.ndarray]:
    path = path or WEIGHTS_PATH
    return np.load(path) if os.path.exists(path) else None


def save_policy(table: np.ndarray, rule_set_hash: int, path: Optional[str] = None) -> None:
    np.savez(path or POLICY_PATH, table=np.asarray(table, dtype=np.uint32), hash=np.int64(rule_set_hash))


def load_policy(rule_set_hash: int, path: Optional[str] = None) -> Optional[np.ndarray]:
    """The stored table if it was derived for the rule set with this hash."""
    path = path or POLICY_PATH
    if not os.path.exists(path):
        return None
    z = np.load(path)
    return z["table"] if int(z["hash"]) == int(rule_set_hash) else None


def as_on_listed(table: np.ndarray) -> np.ndarray:
    """The table with every list-valid entry set to AS (the network still
    decides the steps on bitmap-only sets)."""
    e = np.arange(POLICY_BUCKETS * 2 * 256)
    choice = ((table[e >> 4] >> ((e & 15) * 2).astype(np.uint32)) & 3).astype(np.int64)
    choice[((e >> 8) & 1) == 1] = 2
    return pack_policy(choice)


def as_rule_table() -> np.ndarray:
    """AS on every listed set, iNFA on every bitmap-only one: the fixed rule
    of the AS mode, as a table."""
    e = np.arange(POLICY_BUCKETS * 2 * 256)
    return pack_policy(np.where(((e >> 8) & 1) == 1, 2, 0))


def calibrate(scratch: E.Scratch, buf: np.ndarray, offsets: np.ndarray, stride: int = 4,
              reps: int = 3) -> Optional[np.ndarray]:
    """Compiles the stored network into a policy table for this rule set and
    its traffic, on a sample of strings drawn apart from the ones to be
    scanned: one logging scan, then the table of least predicted cost over
    the logged steps (policy_from_log).  The predictions are per-step
    latencies; whether they also win for the whole scan is checked by
    timing that table against the same table with AS on every listed set,
    on the sample, and against the AS mode's fixed rule (the fallback); the
    fastest is installed.  Returns the installed table, or None when no
    weights are stored."""
    w = load_weights()
    if w is None:
        return None
    d_buf, d_off = E.DeviceArray(buf), E.DeviceArray(offsets)
    n = offsets.size - 1
    rec = collect(scratch, d_buf, d_off, n, cap=1 << 22, stride=stride)
    learned = policy_from_log(scratch, w, from_log(rec))
    best = None
    for table in (learned, as_on_listed(learned), as_rule_table()):
        scratch.set_policy(table)
        ms = float("inf")
        for _ in range(reps + 1):
            scratch.scan_device(d_buf.ptr, d_off.ptr, n, E.KIND_ENGINE, E.MODE_LEARNED)
            scratch.collect(n)
            ms = min(ms, scratch.last_stats()[0])
        if best is None or ms < best[0]:
            best = (ms, table)
    scratch.set_policy(best[1])
    return best[1]


def apply(scratch: E.Scratch, every: int = 0) -> bool:
    """Loads the stored weights, and the policy table derived from them for
    this rule set when one is stored, into the scratch; False when no
    weights are stored."""
    w = load_weights()
    if w is None:
        return False
    scratch.set_selector(w, every=every)
    t = load_policy(scratch.info()["hash"])
    if t is not None and every == 0:
        scratch.set_policy(t)
    return True
