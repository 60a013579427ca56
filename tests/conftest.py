import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
# the package's Python modules (pyanqi, workload, pipeline, ...)
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "anqi-implementation_amd"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")
