import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libctprj.so")


def pytest_collection_modifyitems(config, items):
    # fail loudly if a gpu test is collected where no GPU exists and the user did not
    # deselect it: the GPU tests are the parity tests proper and have no CPU fallback
    pass


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.build()
    return oracle


def golden_cases():
    import glob
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))


def load_golden(name):
    import numpy as np
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    g = z["geometry"]
    geo = (int(g[0]), int(g[1]), int(g[2]), int(g[3]), float(g[4]), float(g[5]), float(g[6]))
    return geo, {k: z[k] for k in z.files if k != "geometry"}


def rel_l2(a, b):
    import numpy as np
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))
