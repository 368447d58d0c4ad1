"""Test configuration: import paths and the `gpu` marker.

GPU tests (marked `gpu`) call libdmpm.so through its C ABI and compare with
the CPU oracle (oracle/liboracle.so).  CPU tests check the oracle itself
against the golden vectors and the mathematics, and check that the HIP
library loads and exports its symbols.
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "shape-morphing-binder_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs libdmpm.so kernels)")
