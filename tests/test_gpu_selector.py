"""Learned selector on the GPU: training-kernel gradients against the FP64
numpy oracle, logging mode, and what the trained network gains."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from admatch_amd import engine as E  # noqa: E402
from admatch_amd import selector  # noqa: E402

import oracle  # noqa: E402
from oracle import mlp_ref  # noqa: E402


def random_problem(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (n, 16)).astype(np.float32)
    y = rng.normal(10, 2, (n, 3)).astype(np.float32)
    mask = (rng.random((n, 3)) < 0.8).astype(np.float32)
    w = selector.init_weights(seed)
    w[-3:] = 10.0
    return x, y, mask, w


@pytest.mark.parametrize("n", [1, 33, 1000])
def test_gradient_matches_fp64_oracle(n):
    x, y, mask, w = random_problem(n, n)
    loss, g = selector.gpu_grad(x, y, mask, w)
    rl, rg = mlp_ref.loss_and_grad(x, y, mask, w)
    assert abs(loss - rl) <= 1e-4 * max(1.0, abs(rl))
    # fp32 accumulation (atomics in any order) vs fp64: relative to the norm
    err = np.abs(g - rg).max() / max(1e-6, np.abs(rg).max())
    assert err < 2e-4, err


def test_gradient_matches_finite_differences():
    x, y, mask, w = random_problem(64, 5)
    _, rg = mlp_ref.loss_and_grad(x, y, mask, w)
    rng = np.random.default_rng(0)
    for i in rng.choice(w.size, 12, replace=False):
        e = np.zeros_like(w, dtype=np.float64)
        e[i] = 1e-5
        lp, _ = mlp_ref.loss_and_grad(x, y, mask, w + e)
        lm, _ = mlp_ref.loss_and_grad(x, y, mask, w - e)
        assert abs((lp - lm) / 2e-5 - rg[i]) < 1e-4 * max(1.0, abs(rg[i]))


def test_forward_matches_oracle():
    x, _, _, w = random_problem(500, 9)
    out = selector.gpu_forward(x, w)
    ref = mlp_ref.forward(x, w)
    assert np.abs(out - ref).max() < 1e-3 * max(1.0, np.abs(ref).max())


def test_training_reduces_loss():
    rng = np.random.default_rng(3)
    x = rng.normal(0, 1, (4096, 16)).astype(np.float32)
    # a learnable target: per-method linear costs of two features
    y = np.stack([10 + x[:, 0], 10 - x[:, 0] + 0.5 * x[:, 2], 9 + x[:, 1]], 1).astype(np.float32)
    ds = selector.Dataset(x=x, y=y, mask=np.ones_like(y))
    w, hist = selector.train(ds, epochs=15, batch=256, lr=3e-3, seed=1)
    assert hist[-1] < 0.1 * hist[0], hist
    pred = selector.gpu_forward(x, w)  # weights take raw features (normalisation folded)
    assert np.mean(pred.argmin(1) == y.argmin(1)) > 0.9


@pytest.mark.parametrize("seed", [3, 4])
def test_policy_table_matches_network(workloads, seed):
    # random weights give varied decisions; the table must hold, for every
    # (bucket, list validity, byte), the FP64 network's choice at the entry's
    # step context (a near tie may go either way in FP32)
    w, path = workloads("medium", 64)
    s = E.Scratch(E.AD_MATCHER, path).alloc_scratch()
    X, allow = selector.policy_inputs(s)
    wts = selector.init_weights(seed)
    wts[-3:] = 0.0
    # output biases that centre each method's prediction, so all three win somewhere
    wts[-3:] = -np.median(mlp_ref.forward(X, wts), axis=0).astype(np.float32)
    s.set_selector(wts)
    e = np.arange(selector.POLICY_BUCKETS * 2 * 256)
    got = selector.policy_lookup(s.policy_table(), np.where(e >> 9 == 0, 0, (3 << np.maximum((e >> 9) - 1, 0)) - 1),
                                 (e >> 8) & 1, e & 255)
    pred = mlp_ref.forward(X, wts)
    pred[~allow, 2] = np.inf
    exp = pred.argmin(1)
    assert set(np.unique(exp).tolist()) == {0, 1, 2}
    bad = np.nonzero(got != exp)[0]
    gap = np.abs(pred[bad, got[bad]] - pred[bad, exp[bad]])
    assert (gap < 1e-3).all(), (bad[:10], gap[:10])
    assert bad.size <= 0.01 * e.size
    # scans that follow this mixed table (a lookup on every step) return the
    # oracle's results under both storages
    buf, off = w.inputs.buf, w.inputs.offsets
    exp_ids = oracle.OracleNFA(path).batch_scan(buf, off, nthreads=4)
    d_buf, d_off = E.DeviceArray(buf), E.DeviceArray(off)
    for store in (0, 1):
        s.set_variant(2, 256, 64, store)
        s.scan_device(d_buf.ptr, d_off.ptr, off.size - 1, E.KIND_ENGINE, E.MODE_LEARNED)
        assert E.unpack(*s.collect(off.size - 1)) == exp_ids, store
    s.close()


def test_log_mode_and_learned_gain(workloads):
    w, path = workloads("large", 1024)
    buf, off = w.inputs.buf, w.inputs.offsets
    n = off.size - 1
    s = E.Scratch(E.AD_MATCHER, path).alloc_scratch()
    d_buf, d_off = E.DeviceArray(buf), E.DeviceArray(off)
    s.scan_device(d_buf.ptr, d_off.ptr, n, E.KIND_ENGINE, E.MODE_HEURISTIC)
    ref = s.collect(n)
    # logging runs all three methods per sampled step and keeps results
    rec = selector.collect(s, d_buf, d_off, n, cap=1 << 18, stride=8)
    assert rec.shape[0] > 1000 and rec.shape[1] == 22
    s.set_log(1 << 18, 8)
    s.scan_device(d_buf.ptr, d_off.ptr, n, E.KIND_ENGINE, E.MODE_LOG)
    got = s.collect(n)
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    s.set_log(0, 8)
    ds = selector.from_log(rec)
    assert (ds.mask[:, 0] == 1).all() and (ds.mask[:, 1] == 1).all()
    assert ((ds.c >= 0) & (ds.c < 256)).all() and (ds.n_expl >= 0).all()
    # the logged total count includes the implicitly held states
    assert (np.rint(np.exp2(ds.x[:, 0]) - 1) >= ds.n_expl).all()
    half = ds.n // 2
    tr = selector.Dataset(ds.x[:half], ds.y[:half], ds.mask[:half], ds.c[:half], ds.n_expl[:half])
    te = selector.Dataset(ds.x[half:], ds.y[half:], ds.mask[half:], ds.c[half:], ds.n_expl[half:])
    wts, hist = selector.train(tr, epochs=20, batch=512, lr=2e-3, seed=0)
    assert hist[-1] < hist[0]
    pred = selector.gpu_forward(te.x, wts)
    learned = selector.regret(te, selector.choose(pred, te.x[:, 10] > 0))
    heur = selector.regret(te, selector.heuristic_choice(te, s.info()["words"]))
    # the network's choices cost fewer measured cycles than the heuristic's
    assert learned["picked_cycles"] < heur["picked_cycles"], (learned, heur)
    # so do those of the policy table the engine reads (least predicted cost
    # per entry over the training steps)
    table = selector.policy_from_log(s, wts, tr)
    s.set_policy(table)
    assert np.array_equal(s.policy_table(), table)
    tab = selector.regret(te, selector.policy_lookup(table, te.n_expl, te.x[:, 10] > 0, te.c))
    assert tab["picked_cycles"] < heur["picked_cycles"], (tab, heur)
    # and the learned engine returns the same rule ids (table from the log,
    # then the network in the scan every 4 steps)
    for every in (0, 4):
        if every:
            s.set_selector(wts, every=every)
        s.scan_device(d_buf.ptr, d_off.ptr, n, E.KIND_ENGINE, E.MODE_LEARNED)
        got = s.collect(n)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    exp = oracle.OracleNFA(path).batch_scan(buf, off, nthreads=8)
    assert E.unpack(*ref) == exp
    s.close()
