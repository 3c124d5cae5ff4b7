"""pytest configuration: the `gpu` marker and import paths.

GPU tests (marked `gpu`) call the HIP library through its C ABI and compare
with the CPU oracle in ``oracle/``; CPU tests pin the oracle itself and check
the host logic and the library's exported symbols without touching a GPU.
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X); calls the HIP library")
