"""The C ABI from a C++ caller: tests/native/sweep_c_abi.cpp includes only
include/qss.h, links libqss.so and sweeps a 32x32, 20-layer wake (the way
the reference's host code would call it).  Its residuals and Ex history are
bitwise those of the same sweep through the Python mirror, which the other
GPU tests hold to the oracle.  The binary is built by
__graft_entry__.build()."""
import os
import subprocess
import sys

import numpy as np
import pytest

from tertia_amd.params import Bunch, Params

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "native", "sweep_c_abi")


def test_cpp_caller_matches_python_mirror(tmp_path):
    if not os.path.exists(BIN):
        raise FileNotFoundError("%s is missing: run __graft_entry__.build()" % BIN)
    out = tmp_path / "ex.bin"
    print("running", BIN, flush=True)
    r = subprocess.run([BIN, str(out)], capture_output=True, text=True, timeout=120)
    sys.stdout.write(r.stdout)
    assert r.returncode == 0, r.stderr
    p = Params(ny=32, nz=32, ly=6.0, lz=6.0, hx=0.05, nslices=20, n_iter=2, dens=1.0,
               jitter=1e-3, seed=12345,
               bunches=[Bunch(nb=2.0, sigma_r=0.5, sigma_xi=0.3, xi0=0.5, polarity=-1.0)])
    data = np.fromfile(out, dtype=np.float64)
    res_c = data[:p.nslices]
    ex_c = data[p.nslices:].reshape(p.nslices, p.nz, p.ny)
    from tertia_amd.mesh import Mesh
    m = Mesh(p, history=("Ex",))
    try:
        res = m.MoveAllSplitLayers()
        ex = np.stack([m.history("Ex", s) for s in range(p.nslices)])
    finally:
        m.close()
    assert np.array_equal(res_c, res)
    assert np.array_equal(ex_c, ex)
    assert np.abs(ex).max() > 0
