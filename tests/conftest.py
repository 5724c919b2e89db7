import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")


def load_cases():
    """Golden vectors from the reference build: {case: {field: array}}."""
    z = np.load(os.path.join(GOLDEN, "ref_cases.npz"), allow_pickle=False)
    cases = {}
    for k in z.files:
        name, field = k.split("/", 1)
        cases.setdefault(name, {})[field] = z[k]
    return cases


@pytest.fixture(scope="session")
def ref_cases():
    return load_cases()


@pytest.fixture(scope="session")
def parse_fixture():
    z = np.load(os.path.join(GOLDEN, "parse_descriptors.npz"), allow_pickle=False)
    out = {}
    for k in z.files:
        name, field = k.split("/", 1)
        out.setdefault(name, {})[field] = z[k]
    return out


def model_path(rel):
    return os.path.join(ROOT, str(rel))


def nbr_sets(counts, entries):
    """per-atom sets of (j, ic0, ic1, ic2) from concatenated rows"""
    out, o = [], 0
    for c in counts:
        out.append(set(map(tuple, np.asarray(entries[o:o + c]).tolist())))
        o += c
    return out
