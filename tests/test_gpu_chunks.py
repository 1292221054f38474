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


def _al(n):
    return (n + 255) // 256 * 256


def test_unpack_refuses_a_foreign_or_damaged_state():
    """qss_unpack_layer checks the buffer's size, its header (format, grid,
    tiling) and the particle cells and tile starts before it touches the
    context: every refusal is QSS_E_ARG and leaves the context able to take
    the good state and continue bitwise."""
    from tertia_amd._lib import QssError
    from tertia_amd.mesh import Mesh
    p = small(ny=32, nz=32, nslices=12)
    a = Mesh(p, history=HIST)
    other = Mesh(small(ny=48, nz=32, nslices=12), history=())
    c = Mesh(p, history=HIST)
    try:
        a.sweep_range(0, 5)
        st = a.pack_layer()
        a.sweep_range(5, 7)
        ref = {f: [a.history(f, s) for s in range(5, 12)] for f in HIST}
        np_ = 4 * 32 * 32
        o_cell = 256 + 5 * _al(np_ * 8)
        o_ts = o_cell + _al(np_ * 4) + _al(np_ * 8)
        bad = []
        b = st.copy(); b[3] ^= 0x20; bad.append(("magic", b))
        b = st.copy(); b.view(np.int32)[o_cell // 4 + 17] = -3; bad.append(("cell", b))
        # 32x32 cells = 2x2 tiles: tile starts 0..4
        b = st.copy(); b.view(np.int32)[o_ts // 4 + 2] = 1 << 30; bad.append(("tile start", b))
        b = st.copy(); b.view(np.int32)[o_ts // 4 + 1] += 1; bad.append(("tile boundary", b))
        bad.append(("truncated", st[:-8].copy()))
        missed = []
        for what, b in bad:
            try:
                c.unpack_layer(b)
                missed.append(what)
            except QssError as e:
                print("refused (%s): %s" % (what, e), flush=True)
                assert e.code == -1, what
        assert not missed, missed
        with pytest.raises(QssError) as e:
            other.unpack_layer(st)
        assert e.value.code == -1
        c.unpack_layer(st)
        c.sweep_range(5, 7)
        for f in HIST:
            for i, s in enumerate(range(5, 12)):
                assert np.array_equal(c.history(f, s), ref[f][i]), (f, s)
    finally:
        a.close()
        other.close()
        c.close()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("ny", [32, 24])
def test_reference_record_round_trip(ny):
    """GPU -> the reference's layer record -> GPU: the continuation rebuilds
    the P-layer spectra from the planes (as the reference's Guess does), so
    it equals the unbroken sweep to rounding.  32x32: tiled particles,
    24x32: the global path."""
    from tertia_amd.mesh import Mesh
    p = small(ny=ny, nz=32, nslices=20)
    k = 8
    full = Mesh(p, history=HIST)
    a = Mesh(p, history=HIST)
    b = Mesh(p, history=HIST)
    try:
        full.MoveAllSplitLayers()
        a.sweep_range(0, k)
        rec = a.PackLayer(p.nslices - k)
        assert rec.size == 12 * ny * 32 + 12 * 4 * ny * 32
        b.UnpackLayer(rec)
        b.sweep_range(k, p.nslices - k)
        for f in HIST:
            G = np.stack([b.history(f, s) for s in range(k, p.nslices)])
            R = np.stack([full.history(f, s) for s in range(k, p.nslices)])
            e = _rel(G, R)
            print("record round trip", ny, f, "%.1e" % e, flush=True)
            assert e <= 1e-12, (f, e)
    finally:
        full.close()
        a.close()
        b.close()


@pytest.mark.parametrize("direction", ["oracle_to_gpu", "gpu_to_oracle"])
def test_reference_record_between_oracle_and_gpu(oracle_mod, direction):
    """A record written by one side after layer k is read by the other, which
    sweeps on; both continuations match the oracle's unbroken sweep (rel-L2
    1e-9, the whole-sweep parity bar)."""
    from tertia_amd.mesh import Mesh
    from tertia_amd.params import Bunch
    p = small(ny=32, nz=32, nslices=20,
              bunches=[Bunch(nb=2.0, sigma_r=0.5, sigma_xi=0.3, xi0=0.5)])
    k = 9
    full = oracle_mod.Oracle(p, history=HIST)
    full.sweep()
    if direction == "oracle_to_gpu":
        src = oracle_mod.Oracle(p, history=())
        src.sweep_range(0, k)
        rec = src.pack_layer_record(p.nslices - k)
        src.close()
        dst = Mesh(p, history=HIST)
        dst.UnpackLayer(rec)
    else:
        src = Mesh(p, history=())
        src.sweep_range(0, k)
        rec = src.PackLayer(p.nslices - k)
        src.close()
        dst = oracle_mod.Oracle(p, history=HIST)
        dst.unpack_layer_record(rec)
    try:
        dst.sweep_range(k, p.nslices - k)
        for f in HIST:
            G = np.stack([dst.history(f, s) for s in range(k, p.nslices)])
            R = np.stack([full.history(f, s) for s in range(k, p.nslices)])
            e = _rel(G, R)
            print(direction, f, "%.1e" % e, flush=True)
            assert e <= 1e-9, (f, e)
    finally:
        dst.close()
        full.close()


def test_reference_record_refuses_other_particles():
    """UnpackLayer of a record whose particles the sweep cannot carry (another
    species' q2m, a weight that is not the plasma's, a cell outside the grid,
    a count that is not ppc * Ny * Nz) is refused with QSS_E_ARG."""
    from tertia_amd._lib import QssError
    from tertia_amd.mesh import Mesh
    p = small(ny=16, nz=16, nslices=6)
    m = Mesh(p, history=())
    try:
        m.sweep_range(0, 2)
        rec = m.PackLayer(p.nslices - 2)
        nc, np_ = 16 * 16, 4 * 16 * 16
        for row, val in ((6, 1.0 / 1836), (7, 0.5), (9, 16.0), (11, 1.0)):
            r = rec.copy()
            r[12 * nc + row * np_ + 3] = val
            with pytest.raises(QssError):
                m.UnpackLayer(r)
        with pytest.raises(QssError):
            m.UnpackLayer(rec[:-12])
        m.UnpackLayer(rec)
    finally:
        m.close()
