import os
import sys

# The CPU oracle parallelises with OpenMP.  The test cases are small, so one thread per process
# is fastest, and several test processes with a full team each oversubscribe the cores (idle
# OpenMP threads spin).  A value set by the environment (e.g. the GPU host's share of cores, for
# the full-size comparisons) is kept.  This must happen before the oracle library is loaded.
os.environ.setdefault("OMP_NUM_THREADS", "1")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD MI355X (gfx950) GPU")
