import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

REF_DATA = os.path.join(ROOT, "tests", "golden", "ref_data")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    # the oracle (checker) is plain C; build it if this checkout has no .so
    lib = os.path.join(ROOT, "oracle", "liby3o.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                       capture_output=True)


def load_columns(name, nheader=0):
    import numpy as np
    rows = []
    with open(os.path.join(REF_DATA, name)) as f:
        for i, line in enumerate(f):
            if i < nheader or not line.strip() or line.lstrip().startswith("#"):
                continue
            rows.append([float(v) for v in line.split()])
    return np.array(rows)


@pytest.fixture(scope="session")
def runtime():
    import y3_cluster_cpp_amd as y3
    rt = y3.Runtime(0)
    yield rt
    rt.close()
