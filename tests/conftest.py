import os
import sys
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests", "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a - b))


def read_su(path):
    b = open(path, "rb").read()
    ns = int(np.frombuffer(b[114:116], np.uint16)[0])
    tr = 240 + 4 * ns
    n = len(b) // tr
    return np.stack([np.frombuffer(b[k * tr + 240:(k + 1) * tr], np.float32) for k in range(n)])


@pytest.fixture(scope="session")
def examples_dir():
    """The reference's example directory rebuilt from tests/golden/examples."""
    import cases
    td = tempfile.mkdtemp(prefix="sofi_examples_")
    cases.materialize_examples(cases.EXAMPLES, td)
    return td


class chdir:
    def __init__(self, d):
        self.d = d

    def __enter__(self):
        self.old = os.getcwd()
        os.chdir(self.d)

    def __exit__(self, *a):
        os.chdir(self.old)
