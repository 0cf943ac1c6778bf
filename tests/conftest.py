"""Test configuration: import paths and the `gpu` marker.

`-m "not gpu"` runs the oracle pins, host logic and ABI checks on the CPU;
`-m gpu` runs the parity tests of the HIP library against the oracle."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "smoothedparticles.jl_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
