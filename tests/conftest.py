import os
import sys

import pytest

# Under pytest-xdist every worker would start one BLAS thread per core; on an
# 8-core host four workers then slow the dense oracle solves many times over.
# Must run before numpy is imported.
if "PYTEST_XDIST_WORKER" in os.environ:
    for _v in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
        os.environ.setdefault(_v, "2")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (MI355X)")


@pytest.fixture(scope="session")
def lib():
    from gds2para_amd import _capi
    return _capi.load()


@pytest.fixture(scope="session")
def microstrip():
    from gds2para_amd import mesh as M
    return M.Mesh(M.microstrip_spec())


@pytest.fixture(scope="session")
def small_manhattan():
    from gds2para_amd import mesh as M
    return M.Mesh(M.manhattan_spec((24, 20, 10), n_ports=3, nfreq=3, f_range=(1e8, 1e10), seed=11))
