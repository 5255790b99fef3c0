import gzip
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


def golden_nfa(name: str, tmp_dir: str) -> str:
    """Path of a (decompressed) golden .nfa fixture."""
    src = os.path.join(GOLDEN, name)
    if not name.endswith(".gz"):
        return src
    dst = os.path.join(tmp_dir, name[:-3])
    if not os.path.exists(dst):
        with gzip.open(src, "rb") as fi, open(dst, "wb") as fo:
            fo.write(fi.read())
    return dst


def golden_inputs(name: str):
    z = np.load(os.path.join(GOLDEN, name))
    return z["buf"], z["offsets"]


def golden_json(name: str):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def ref_fixture_files(tmp_dir: str, subset: str):
    """Paths of the per-rule .nfa files of data/<subset> (crs or snort),
    extracted once from tests/golden/ref_nfa_fixtures.tar.gz."""
    import tarfile
    root = os.path.join(tmp_dir, "ref_nfa")
    if not os.path.isdir(root):
        with tarfile.open(os.path.join(GOLDEN, "ref_nfa_fixtures.tar.gz")) as t:
            t.extractall(root, filter="data")
    d = os.path.join(root, subset)
    return sorted(os.path.join(d, f) for f in os.listdir(d) if f.endswith(".nfa"))


@pytest.fixture(scope="session")
def gold_dir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("golden"))


@pytest.fixture(scope="session")
def workloads(tmp_path_factory):
    """Synthetic workloads, generated once per session, with .nfa paths."""
    from admatch_amd import synth
    d = str(tmp_path_factory.mktemp("synth"))
    cache = {}

    def get(name, n=None):
        key = (name, n)
        if key not in cache:
            w = synth.make_workload(name, n=n)
            path = os.path.join(d, f"{name}.nfa")
            if not os.path.exists(path):
                w.ruleset.write(path)
            cache[key] = (w, path)
        return cache[key]

    return get


def gpu_available() -> bool:
    try:
        from admatch_amd.engine import Hip
        return Hip.device_count() > 0
    except Exception:
        return False
