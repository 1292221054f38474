"""GPU: batched contexts (several independent wakes advanced by the same
kernel launches) give each wake exactly what a single-wake context gives
(bitwise, with the same transform tile shapes), in FP64 and MIXED32."""
import numpy as np
import pytest

from tertia_amd.params import Bunch, PRECISIONS, parameter_scan, small

pytestmark = pytest.mark.gpu

WAKES = [[Bunch(nb=0.5, sigma_r=0.5, sigma_xi=0.4, xi0=0.6)],
         [Bunch(nb=3.0, sigma_r=0.3, sigma_xi=0.3, xi0=0.5)],
         [Bunch(nb=1.5, sigma_r=0.7, sigma_xi=0.5, xi0=0.8, noise=0.02)]]


def run(p, wake=None):
    from tertia_amd.mesh import Mesh
    m = Mesh(p, history=("Ex", "Psi"))
    try:
        res = m.MoveAllSplitLayers()
        out = []
        for w in range(p.nwake):
            m.select_wake(w)
            out.append({"Ex": np.stack([m.history("Ex", s) for s in range(p.nslices)]),
                        "Psi": np.stack([m.history("Psi", s) for s in range(p.nslices)]),
                        "part": m.particles()})
    finally:
        m.close()
    return res, out


@pytest.mark.parametrize("prec", ["fp64", "mixed32"])
def test_batch_equals_single_wakes(prec):
    pb = small(ny=32, nz=48, nslices=24, precision=PRECISIONS[prec],
               bunches=list(WAKES[0]), wakes=WAKES)
    res_b, out_b = run(pb)
    assert res_b.shape == (len(WAKES), pb.nslices)
    for w, bl in enumerate(WAKES):
        ps = small(ny=32, nz=48, nslices=24, precision=PRECISIONS[prec], bunches=list(bl))
        res_s, out_s = run(ps)
        assert np.array_equal(res_s, res_b[w]), w
        for f in ("Ex", "Psi"):
            assert np.array_equal(out_s[0][f], out_b[w][f]), (w, f)
        for k in ("y", "z", "px", "py", "pz", "id"):
            assert np.array_equal(out_s[0]["part"][k], out_b[w]["part"][k]), (w, k)
    # the wakes differ from each other
    assert not np.array_equal(out_b[0]["Ex"], out_b[1]["Ex"])


def test_parameter_scan_groups_run():
    """The parameter-scan layout: 4 grid sizes x 2 wakes, a few layers."""
    from tertia_amd.mesh import Mesh
    for p in parameter_scan(per_size=2, nslices=6):
        m = Mesh(p, history=())
        try:
            m.autotune(1)
            r = m.MoveAllSplitLayers()
        finally:
            m.close()
        assert r.shape == (2, 6) and np.all(np.isfinite(r))
