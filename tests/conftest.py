import os
import sys

import pytest

# Under pytest-xdist every worker process loads the OpenMP restatement: share
# the cores among the workers (all threads in every worker oversubscribe the
# machine and OpenMP's spinning waits slow it down a hundredfold).
if "PYTEST_XDIST_WORKER_COUNT" in os.environ and "OMP_NUM_THREADS" not in os.environ:
    os.environ["OMP_NUM_THREADS"] = str(max(1, (os.cpu_count() or 1) // int(os.environ["PYTEST_XDIST_WORKER_COUNT"])))

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    config.addinivalue_line("markers", "slow: longer CPU cases")
