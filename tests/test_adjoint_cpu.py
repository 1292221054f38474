"""CPU checks of the adjoint sweep's building blocks, compiled with g++ from
the same headers the GPU kernels use:

* push_adj.h: the hand-written reverse of one particle's push (CIC gather,
  Boris rotation, sub-stepped CIC deposit with the sub-cycle restart) against
  central finite differences in every input; relative error <= 1e-6 of the
  gradient's max entry, on 400 random particles of which > 10 sub-cycle.
* spectral_mode.h: <M x, y> = <x, M^T y> for the Guess and Iterate per-mode
  maps (and their linearity), to 1e-13."""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "tertia_amd", "csrc")
NATIVE = os.path.join(ROOT, "tests", "native")


def _run(src, *args):
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "t")
        subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", CSRC,
                        os.path.join(NATIVE, src), "-o", exe], check=True)
        return subprocess.run([exe, *args], capture_output=True, text=True, timeout=120)


def test_push_adjoint_matches_finite_differences():
    r = _run("push_adj_check.cpp", "400")
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


def test_spectral_transposes():
    r = _run("spectral_adj_check.cpp")
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


def test_autograd_refuses_a_stale_backward():
    """SweepEx.backward refuses when the mesh has run another checkpointed
    sweep (or changed its bins) since the forward it belongs to; the check
    comes before any device work."""
    import types

    import pytest
    import torch

    from tertia_amd.autograd import SweepEx
    mesh = types.SimpleNamespace(checkpoint_generation=3)
    ctx = types.SimpleNamespace(mesh=mesh, generation=2, bins_device=torch.device("cpu"))
    with pytest.raises(RuntimeError, match="another checkpointed sweep"):
        SweepEx.backward(ctx, torch.zeros(1, dtype=torch.float64))
