"""The C++ mio:: call-shape layer (include/mgo_mio.hpp) on the device: the reference's graph example,
a random graph with every feature (fixed step, advance in pieces), a 16-run ParameterStudy::run batch and
osecir::simulate with RKIntegratorCore, each against the oracle; errors surface as mio::Error."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu


def test_cpp_shim_on_device():
    exe = os.path.join(os.path.dirname(__file__), "cpp", "test_mio_shim")
    assert os.path.exists(exe), "tests/cpp/test_mio_shim is built by __graft_entry__.build()"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0 and "ALL PASSED" in r.stdout, r.stdout + r.stderr
