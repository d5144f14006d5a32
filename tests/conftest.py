import glob
import importlib
import importlib.util
import os
import sys

import numpy as np
import pytest

# torch (device buffers and streams of some GPU tests) first, whichever tests
# run: the HIP runtime it brings is then the one the library binds, as in
# bench.py; loaded after the library, torch found no device
try:
    import torch  # noqa: F401
except ImportError:
    pass

ROOT =os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def golden_cases():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def fmm_module():
    """The product package (its directory name starts with a digit)."""
    return importlib.import_module("2dfmm_amd")


def _load_file(name, rel):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(ROOT, rel))
    m = importlib.util.module_from_spec(spec)
    sys.modules[name] = m
    spec.loader.exec_module(m)
    return m


def workloads():
    """Input generators, loaded without the package (which needs the built library)."""
    return _load_file("fmm2d_workloads", os.path.join("2dfmm_amd", "workloads.py"))


def sfile():
    return _load_file("fmm2d_sfile", os.path.join("2dfmm_amd", "sfile.py"))


def rel_err(a, b):
    """(rel_L2, rel_max) of a against reference b, normwise over all targets."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    mb = np.abs(b).max() if b.size else 0.0
    l2 = np.linalg.norm(a - b) / nb if nb > 0 else np.linalg.norm(a - b)
    mx = np.abs(a - b).max() / mb if mb > 0 else (np.abs(a - b).max() if a.size else 0.0)
    return float(l2), float(mx)


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build(ref=True)
    return O
