"""GPU against the FP64 oracle at the full 512x512 x 4 ppc two-bunch workload
(20x20 k_p^-1, 1,048,576 plasma electrons per slice, dxi = 0.01, driver
nb/n0 = 4 at xi = 2.1, witness nb/n0 = 8, sigma_r = 0.1 at xi = 6.6), through
slice 704 (xi = 7.04), past the witness.  The oracle runs the FFT-form
DCT-IV on 16 threads (about two minutes).

Stated per run, every 8th layer (88 layers): relative L2 and max error
(relative to the field's max) of Ex (the longitudinal wake, E_z in the usual
wakefield notation), Ey, Ez, Bx, By, Bz and psi, and the max deviation of the
plasma particle tracks after the last layer.

  * FP64: rel-L2 and max <= 1e-9 per field, tracks <= 1e-6 (both sides are
    FP64 and differ by summation order only);
  * bf16x6 (split-BF16 transforms with six products, FP32 push, FP64
    particle state and residual): Ex and psi within rel-L2 <= 1e-5 of the
    FP64 oracle; the other fields are stated.
"""
import numpy as np
import pytest

from tertia_amd.params import PRECISIONS, two_bunch_512

pytestmark = pytest.mark.gpu

NS = 704
EVERY = 8
FIELDS = ("Ex", "Ey", "Ez", "Bx", "By", "Bz", "Psi")
SLICES = list(range(EVERY - 1, NS, EVERY))


@pytest.fixture(scope="module")
def oracle_run(oracle_mod):
    p = two_bunch_512(nslices=NS)
    o = oracle_mod.Oracle(p, use_fft=True, nthreads=16, history=FIELDS)
    try:
        st, res = o.sweep()
        assert st == 0
        ref = {f: np.stack([o.history(f, s) for s in SLICES]) for f in FIELDS}
        parts = o.particles()
    finally:
        o.close()
    return ref, parts


def gpu_run(precision):
    from tertia_amd.mesh import Mesh
    p = two_bunch_512(nslices=NS, precision=PRECISIONS[precision])
    m = Mesh(p, history=FIELDS)
    try:
        res = m.MoveAllSplitLayers()
        out = {f: np.stack([m.history(f, s) for s in SLICES]) for f in FIELDS}
        parts = m.particles()
    finally:
        m.close()
    return res, out, parts


def compare(out, ref):
    stats = {}
    for f in FIELDS:
        G, R = out[f], ref[f]
        stats[f] = (float(np.linalg.norm(G - R) / np.linalg.norm(R)),
                    float(np.abs(G - R).max() / np.abs(R).max()))
    return stats


def tracks(pg, po, ny, nz):
    og, oo = np.argsort(pg["id"]), np.argsort(po["id"])
    out = {}
    for k, n in (("y", ny), ("z", nz)):   # global positions, periodic
        d = np.abs(pg[k][og] - po[k][oo])
        out[k] = float(np.minimum(d, n - d).max())
    for k in ("px", "py", "pz"):
        out[k] = float(np.abs(pg[k][og] - po[k][oo]).max())
    return out


@pytest.mark.timeout(900)
def test_two_bunch_512_fp64_vs_oracle(oracle_run):
    ref, po = oracle_run
    res, out, pg = gpu_run("fp64")
    assert np.all(np.isfinite(res))
    assert np.abs(ref["Ex"][SLICES.index(655)]).max() > 0.1   # a real wake at the witness
    stats = compare(out, ref)
    tr = tracks(pg, po, 512, 512)
    print("two_bunch_512 fp64 vs oracle rel-L2/max",
          {f: "%.1e/%.1e" % v for f, v in stats.items()},
          "tracks", {k: "%.1e" % v for k, v in tr.items()}, flush=True)
    for f, (l2, mx) in stats.items():
        assert l2 <= 1e-9 and mx <= 1e-9, (f, l2, mx)
    for k, d in tr.items():
        assert d <= 1e-6, (k, d)


@pytest.mark.timeout(900)
def test_two_bunch_512_bf16x6_vs_oracle(oracle_run):
    ref, po = oracle_run
    res, out, pg = gpu_run("bf16x6")
    assert np.all(np.isfinite(res))
    stats = compare(out, ref)
    tr = tracks(pg, po, 512, 512)
    print("two_bunch_512 bf16x6 vs oracle rel-L2/max",
          {f: "%.1e/%.1e" % v for f, v in stats.items()},
          "tracks", {k: "%.1e" % v for k, v in tr.items()}, flush=True)
    assert stats["Ex"][0] <= 1e-5 and stats["Psi"][0] <= 1e-5, stats
