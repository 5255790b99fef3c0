"""The C ABI library loads and exports every symbol include/admatch.h
declares; host-only calls work without a GPU."""
import os
import re
import subprocess

import pytest

from conftest import ROOT, gpu_available, golden_nfa

from admatch_amd import engine as E


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "admatch.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\**(ita_\w+)\s*\(", text, re.M)))


def test_exports_every_declared_symbol():
    syms = declared_symbols()
    assert len(syms) >= 17
    lib = E.lib()
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(E.EXPORTED)


def test_exports_cpp_surface():
    out = subprocess.run(["nm", "-D", "--defined-only", E.LIB_PATH], capture_output=True, text=True).stdout
    dem = subprocess.run(["c++filt"], input=out, capture_output=True, text=True).stdout
    for sig in ["ita_scratch::ita_scratch(unsigned short, char*)", "ad_scratch::ad_scratch(unsigned short, char*)",
                "allocScratch(ita_scratch&)", "allocScratch(ad_scratch&)", "freeScratch(ita_scratch&)",
                "freeScratch(ad_scratch&)", "Scan(ita_scratch&, char*, std::vector<int, std::allocator<int> >*)",
                "Scan(ad_scratch&, char*, std::vector<int, std::allocator<int> >*)",
                "BatchedScan(ita_scratch&, char**, int*, int, std::vector<int, std::allocator<int> >*)",
                "BatchedScan(ad_scratch&, char**, int*, int, std::vector<int, std::allocator<int> >*)"]:
        assert sig in dem, sig


def test_info_without_gpu(gold_dir):
    s = E.AdScratch(0, golden_nfa("crs_g6.nfa.gz", gold_dir))
    info = s.info()
    assert info["states"] == 1997 and info["words"] == 32
    assert info["tko_groups"] > 0 and info["as_entries"] == info["transitions"] - info["wb_transitions"]
    assert s.variant() == (2, 512, 64, 0)
    lds_bitmaps = s.info()["engine_lds_bytes"]
    s.set_variant(4, 1024, 128)
    assert s.variant() == (4, 1024, 128, 0)
    s.set_variant(4, 256, 128, store=1)
    assert s.variant() == (4, 256, 128, 1)
    # hashed storage: two lists of cap + 64 u32 and two tables of 2*cap + 64 u32 per string
    assert s.info()["engine_lds_bytes"] == 4 * (2 * (256 + 64) + 2 * (512 + 64)) * 4
    s.set_variant(2, 512, 64)
    assert s.info()["engine_lds_bytes"] == lds_bitmaps
    with pytest.raises(E.AdmatchError):
        s.set_variant(3, 512, 64)
    with pytest.raises(E.AdmatchError):
        s.set_variant(2, 512, 64, store=2)
    s.close()


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU error path")
def test_alloc_without_gpu_is_an_error(gold_dir):
    s = E.ItaScratch(E.INFA_KERNEL, golden_nfa("crs_rule1.nfa", gold_dir))
    with pytest.raises(E.AdmatchError) as ei:
        s.alloc_scratch()
    assert ei.value.rc == -3
    with pytest.raises(E.AdmatchError) as ei:
        s.batched_scan([b"text/xml"])
    assert ei.value.rc == -4  # scan before allocScratch


def test_tools_help():
    out = subprocess.run([os.path.join(ROOT, "admatch_amd", "bin", "batch"), "-h"], capture_output=True, text=True)
    assert out.returncode == 0 and "-k" in out.stdout
    out = subprocess.run([os.path.join(ROOT, "admatch_amd", "bin", "batch"), "-n", "0"], capture_output=True,
                         text=True)
    assert out.returncode != 0 and "invalid string_count" in out.stderr


def test_shim_load_failure_exits():
    out = subprocess.run([os.path.join(ROOT, "admatch_amd", "bin", "batch"), "-f", "/nonexistent.nfa", "-i",
                          "/dev/null", "-n", "1", "-k", "AD"], capture_output=True, text=True)
    assert out.returncode == 255 and "Error: load NFA file /nonexistent.nfa" in out.stderr
