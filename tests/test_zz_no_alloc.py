"""GPU: the allocation-free entry point (ita_scan_device_noalloc) against the
entry point it varies, ita_scan_device + ita_collect, on the same inputs:
rule ids equal bit for bit, on consecutive calls (the second one smaller,
the next on another handle) that share the library's device array."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from admatch_amd import engine as E  # noqa: E402
from admatch_amd import selector  # noqa: E402


def rules_of(s, counts, starts, states):
    off, ids = s.accept_map()
    out = []
    for c, st in zip(counts.tolist(), starts.tolist()):
        assert c >= 0
        got = states[st:st + c]
        assert (np.diff(got.astype(np.int64)) > 0).all()  # ascending, distinct
        r = set()
        for x in got.tolist():
            r.update(ids[off[x]:off[x + 1]].tolist())
        out.append(sorted(r))
    return out


class NoAllocRun:
    """Device buffers of one call, allocated here by the test (the caller)."""

    def __init__(self, buf, off, states_cap):
        self.n = off.size - 1
        self.d_buf, self.d_off = E.DeviceArray(buf), E.DeviceArray(off)
        self.d_cnt = E.DeviceArray(np.full(max(self.n, 1), -7, np.int32))
        self.d_st = E.DeviceArray(np.zeros(max(self.n, 1), np.int64))
        self.cap = states_cap
        self.d_states = E.DeviceArray(np.zeros(max(states_cap, 1), np.uint32))
        self.d_status = E.DeviceArray(np.full(2, -1, np.int64))

    def scan(self, s, kind, mode):
        s.scan_device_noalloc(self.d_buf.ptr, self.d_off.ptr, self.n, kind, mode, self.d_cnt.ptr, self.d_st.ptr,
                              self.d_states.ptr, self.cap, self.d_status.ptr)
        E.Hip.check(E.Hip.rt().hipDeviceSynchronize(), "hipDeviceSynchronize")
        cnt = self.d_cnt.to_host(np.zeros(self.n, np.int32))
        st = self.d_st.to_host(np.zeros(self.n, np.int64))
        states = self.d_states.to_host(np.zeros(self.cap, np.uint32))
        status = self.d_status.to_host(np.zeros(2, np.int64))
        return cnt, st, states, status


def reference(s, buf, off, kind, mode):
    d_buf, d_off = E.DeviceArray(buf), E.DeviceArray(off)
    s.scan_device(d_buf.ptr, d_off.ptr, off.size - 1, kind, mode)
    return E.unpack(*s.collect(off.size - 1))


def check_call(s, buf, off, kind, mode, what):
    print(f"[case] no-alloc {what} kind={kind} mode={E.MODE_NAMES[mode]} n={off.size - 1}", flush=True)
    exp = reference(s, buf, off, kind, mode)
    run = NoAllocRun(buf, off, states_cap=8 * (off.size - 1) + 64)
    cnt, st, states, status = run.scan(s, kind, mode)
    assert status[0] == 0 and status[1] == int(cnt.sum())
    assert rules_of(s, cnt, st, states) == exp
    return exp, int(status[1])


@pytest.mark.parametrize("mode", [E.MODE_LEARNED, E.MODE_AS, E.MODE_HEURISTIC, E.MODE_TKO])
def test_noalloc_equals_scan_device(workloads, mode):
    w, path = workloads("medium", 2048)
    s = E.Scratch(E.AD_MATCHER, path).alloc_scratch()
    s.set_variant(2, 256, 64, 1)  # the existing path scans with hashed storage
    if mode == E.MODE_LEARNED:
        s.set_selector(selector.load_weights())
    buf, off = w.inputs.buf, w.inputs.offsets
    check_call(s, buf, off, E.KIND_ENGINE, mode, "medium 2048")
    # the second call on the same device array: fewer, different strings
    sub = np.arange(1500, 1800)
    b2, o2 = E.pack_strings([w.inputs.string(int(i)) for i in sub])
    check_call(s, b2, o2, E.KIND_ENGINE, mode, "medium 300")
    s.close()


def test_noalloc_plain_and_another_handle(workloads):
    wm, pm = workloads("medium", 2048)
    ws, ps = workloads("small", 1024)
    a = E.Scratch(E.AD_MATCHER, pm).alloc_scratch()
    b = E.Scratch(E.INFA_KERNEL, ps).alloc_scratch()
    check_call(a, wm.inputs.buf, wm.inputs.offsets, E.KIND_ENGINE, E.MODE_ADMATCH, "medium")
    b2, o2 = E.pack_strings([ws.inputs.string(i) for i in range(700)])
    check_call(b, b2, o2, E.KIND_PLAIN, E.MODE_INFA, "small 700 (other handle)")
    check_call(b, b2, o2, E.KIND_ENGINE, E.MODE_LEARNED, "small 700 untrained")
    a.close()
    b.close()


def test_noalloc_small_output_and_limits(workloads):
    w, path = workloads("small", 1024)
    s = E.Scratch(E.AD_MATCHER, path).alloc_scratch()
    buf, off = w.inputs.buf, w.inputs.offsets
    _, total = check_call(s, buf, off, E.KIND_ENGINE, E.MODE_HEURISTIC, "small")
    assert total > 4
    # an output buffer too small for every list: reported, nothing written past it
    run = NoAllocRun(buf, off, states_cap=4)
    cnt, st, states, status = run.scan(s, E.KIND_ENGINE, E.MODE_HEURISTIC)
    assert status[0] == 1 and status[1] == total and (cnt < 0).any()
    # empty batch
    run0 = NoAllocRun(np.zeros(1, np.uint8), np.zeros(1, np.int64), 4)
    _, _, _, status0 = run0.scan(s, E.KIND_ENGINE, E.MODE_AS)
    assert status0.tolist() == [0, 0]
    # more strings than the array's order can hold: rejected before any launch
    big = (1 << 20) + 1
    with pytest.raises(E.AdmatchError) as ei:
        s.scan_device_noalloc(run.d_buf.ptr, run.d_off.ptr, big, E.KIND_ENGINE, E.MODE_AS, run.d_cnt.ptr,
                              run.d_st.ptr, run.d_states.ptr, 4, run.d_status.ptr)
    assert ei.value.rc == -6 and "at most" in str(ei.value)
    # then a normal call still works
    check_call(s, buf, off, E.KIND_ENGINE, E.MODE_AS, "small after rejection")
    s.close()
