"""GPU: sweeps replayed from a captured hipGraph (pairs of layer steps, the
layer index read from device memory) equal the eagerly launched sweeps
bitwise: every replay is checked (every layer of the history), with eager
layers and host synchronisations between replays (odd sweep ranges force a
parity change and eager layers)."""
import numpy as np
import pytest

from tertia_amd.params import Bunch, PRECISIONS, linear_wake_64, small

pytestmark = pytest.mark.gpu

FIELDS = ("Ex", "Ez", "By", "Psi", "Rho")


def sweep(p, graph, pieces=None):
    from tertia_amd.mesh import Mesh
    m = Mesh(p, history=FIELDS)
    try:
        m.enable_graph(graph)
        if pieces is None:
            res = m.MoveAllSplitLayers()
        else:
            m.reset()
            res, s = [], 0
            for n in pieces:
                res.append(m.sweep_range(s, n))
                s += n
            res = np.concatenate(res, axis=-1)
        out = []
        for w in range(p.nwake):
            m.select_wake(w)
            out.append({f: np.stack([m.history(f, s) for s in range(p.nslices)])
                        for f in FIELDS})
            out[-1]["part"] = m.particles()
    finally:
        m.close()
    return res, out


def same(a, b):
    ra, oa = a
    rb, ob = b
    assert np.array_equal(ra, rb)
    for x, y in zip(oa, ob):
        for f in FIELDS:
            assert np.array_equal(x[f], y[f]), f
        for k in ("y", "z", "px", "py", "pz", "id"):
            assert np.array_equal(x["part"][k], y["part"][k]), k


@pytest.mark.parametrize("prec", ["fp64", "mixed32"])
def test_graph_sweep_is_the_eager_sweep(prec):
    p = small(nslices=21, precision=PRECISIONS[prec])
    same(sweep(p, False), sweep(p, True))


def test_graph_with_odd_ranges_and_linear_wake():
    p = linear_wake_64(nslices=40)
    eager = sweep(p, False)
    same(eager, sweep(p, True))
    # ranges of odd lengths: parity switches, eager layers between replays
    same(eager, sweep(p, True, pieces=[3, 7, 1, 12, 5, 12]))


def test_graph_batched():
    wakes = [[Bunch(nb=0.5, sigma_r=0.5, sigma_xi=0.4, xi0=0.6)],
             [Bunch(nb=2.0, sigma_r=0.3, sigma_xi=0.3, xi0=0.5)]]
    p = small(ny=32, nz=32, nslices=15, bunches=list(wakes[0]), wakes=wakes)
    same(sweep(p, False), sweep(p, True))
