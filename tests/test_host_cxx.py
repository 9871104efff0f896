"""The C++ host layer (chroma_amd/host): the propagator driver is built against the library and runs
without a GPU as far as its argument handling goes."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "chroma_amd", "host", "mi355x_prop")


def test_driver_built_and_usage():
    if not os.path.exists(DRIVER):
        pytest.fail("chroma_amd/host/mi355x_prop is not built (python -m chroma_amd.build)")
    p = subprocess.run([DRIVER, "--help"], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0 and "usage: mi355x_prop" in p.stdout
    p = subprocess.run([DRIVER, "--cfg", "nonsense"], capture_output=True, text=True, timeout=60)
    assert p.returncode == 2 and "unknown --cfg" in p.stderr
