"""A sweep split into xi chunks, each chunk swept by its own context that
continues from the packed state of the chunk in front (the reference's
pipeline over ranks: PackLayer / SendLayer / ReceiveLayer, para.cpp:94-412),
gives the layers of one unbroken sweep bitwise: tiled and global particle
paths, split-BF16 transforms (whose split-operand records unpacking drops).
"""
import numpy as np
import pytest

from tertia_amd.params import QSS_PREC_FP64, QSS_PREC_MIXED_BF16X6, small

pytestmark = pytest.mark.gpu

HIST = ("Ex", "Ey", "Ez", "Bx", "By", "Bz", "Psi")


@pytest.mark.parametrize("ny,cuts,prec", [(32, (7, 15), QSS_PREC_FP64),
                                          (24, (10,), QSS_PREC_FP64),
                                          (32, (12,), QSS_PREC_MIXED_BF16X6)])
def test_chunked_sweep_equals_unbroken(ny, cuts, prec):
    from tertia_amd.mesh import Mesh
    p = small(ny=ny, nz=32, nslices=24, precision=prec)
    full = Mesh(p, history=HIST)
    try:
        full.MoveAllSplitLayers()
        ref = {f: np.stack([full.history(f, s) for s in range(p.nslices)]) for f in HIST}
    finally:
        full.close()
    bounds = [0, *cuts, p.nslices]
    state = None
    for a, b in zip(bounds[:-1], bounds[1:]):
        m = Mesh(p, history=HIST)
        try:
            if state is None:
                m.reset()
            else:
                m.unpack_layer(state)
            m.sweep_range(a, b - a)
            for f in HIST:
                for s in range(a, b):
                    assert np.array_equal(m.history(f, s), ref[f][s]), (f, s)
            state = m.pack_layer()
        finally:
            m.close()
    assert state.nbytes > 0
