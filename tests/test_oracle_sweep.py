"""Whole-sweep invariant of the oracle, independent of any GPU result.

The corrector takes rho from continuity, d(rho - Jx)/dxi = div J_perp, with
the trapezoid rule over a layer, and solves laplacian(Ex) = div(J_perp) and
-laplacian(psi) = rho - Jx per mode (movefieldshydrolinlayersplit.cpp:
1373-1478; psi is not in the reference).  Together they fix, for every pair
of consecutive layers and every cell,

    (psi[s] - psi[s-1]) / hx = (Ex[s] + Ex[s-1]) / 2,

exactly up to rounding.  A dropped term or a wrong sign in the continuity
update, the Ex solve or the psi solve breaks it at O(1)."""
import numpy as np
import pytest

from tertia_amd.params import Bunch, small


@pytest.mark.parametrize("shape", [0, 1])
def test_psi_ex_identity(oracle_mod, shape):
    p = small(ny=32, nz=32, nslices=30, shape=shape,
              bunches=[Bunch(nb=2.0, sigma_r=0.5, sigma_xi=0.4, xi0=0.8)])
    o = oracle_mod.Oracle(p, history=("Ex", "Psi"))
    try:
        st, _ = o.sweep()
        assert st == 0
        ex = np.stack([o.history("Ex", s) for s in range(p.nslices)])
        psi = np.stack([o.history("Psi", s) for s in range(p.nslices)])
    finally:
        o.close()
    d = (psi[1:] - psi[:-1]) / p.hx
    tr = 0.5 * (ex[1:] + ex[:-1])
    assert np.abs(tr).max() > 0.05            # a real wake, not zeros
    assert np.abs(d - tr).max() <= 1e-12 * np.abs(tr).max()
