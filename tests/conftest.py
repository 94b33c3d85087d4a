"""Test setup: the repository root and oracle/ on sys.path, and the `gpu` marker (tests that need a HIP device;
run with -m gpu on a GPU machine, deselected with -m "not gpu")."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built libtdp.so")
