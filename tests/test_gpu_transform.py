"""GPU parity: the MFMA half-integer transform (qss_transform2d) against the
oracle and against the scipy fixtures of the reference's recipe inputs.

Tolerance: FP64 GEMM vs FP64 O(N^2)-per-row sums, max |diff| / max |ref|
<= 1e-13 (both are dot products of N terms with ~1 ulp table entries)."""
import os

import numpy as np
import pytest

from tests.helpers import max_rel
from tertia_amd.params import small

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "transforms.npz")


@pytest.fixture(scope="module")
def mesh_factory():
    from tertia_amd.mesh import Mesh
    made = {}

    def get(nz, ny):
        if (nz, ny) not in made:
            made[(nz, ny)] = Mesh(small(ny=ny, nz=nz, nslices=1), history=())
        return made[(nz, ny)]
    yield get
    for m in made.values():
        m.close()


@pytest.mark.parametrize("nz,ny", [(8, 16), (16, 8), (32, 32), (12, 20)])
def test_transform_recipe_fixtures(mesh_factory, nz, ny):
    g = np.load(GOLD)
    m = mesh_factory(nz, ny)
    x = g["sin_%dx%d_in" % (nz, ny)]
    for kz in (0, 1):
        for ky in (0, 1):
            got = m.transform2d(x, kz, ky)
            ref = g["sin_%dx%d_k%d%d" % (nz, ny, kz, ky)]
            assert max_rel(got, ref) < 1e-13, (kz, ky)


def test_transform_linear_recipe(mesh_factory):
    g = np.load(GOLD)
    m = mesh_factory(16, 16)
    assert max_rel(m.transform2d(g["lin_a_in"], 1, 0), g["lin_a_k10"]) < 1e-13
    assert max_rel(m.transform2d(g["lin_b_in"], 0, 1), g["lin_b_k01"]) < 1e-13


@pytest.mark.parametrize("nz,ny", [(64, 64), (40, 72), (128, 96), (256, 256), (512, 512)])
def test_transform_vs_oracle(mesh_factory, oracle_mod, nz, ny):
    rng = np.random.default_rng(nz * 1000 + ny)
    m = mesh_factory(nz, ny)
    x = rng.standard_normal((nz, ny))
    for kz in (0, 1):
        for ky in (0, 1):
            got = m.transform2d(x, kz, ky)
            ref = oracle_mod.transform2d(x, kz, ky, use_fft=False, nthreads=4)
            assert max_rel(got, ref) < 1e-13, (kz, ky)


def test_inverse_roundtrip(mesh_factory):
    """T.T = 2N.I per axis: forward then scaled forward is the identity."""
    nz, ny = 128, 96
    m = mesh_factory(nz, ny)
    x = np.random.default_rng(5).standard_normal((nz, ny))
    for kz in (0, 1):
        for ky in (0, 1):
            y = m.transform2d(x, kz, ky)
            back = m.transform2d(y, kz, ky, scale=1.0 / (4.0 * nz * ny))
            assert max_rel(back, x) < 1e-13


def test_every_tile_shape(mesh_factory, oracle_mod):
    """Every compiled GEMM tile shape gives the same transform (to 1e-13),
    including partial tiles (40 x 72) and the full 512 x 512 plane."""
    from tertia_amd.mesh import Mesh
    for nz, ny in ((40, 72), (512, 512)):
        m = mesh_factory(nz, ny)
        x = np.random.default_rng(11).standard_normal((nz, ny))
        ref = oracle_mod.transform2d(x, 1, 0, use_fft=False, nthreads=4)
        method = m.transform_method
        m.transform_method = 0                      # the dense tables' kernels
        for v in range(Mesh.gemm_variants()):
            m.set_gemm_variant(v)
            assert max_rel(m.transform2d(x, 1, 0), ref) < 1e-13, (nz, ny, v)
        m.set_gemm_variant(-1)
        choice = m.autotune(2)
        assert all(0 <= c < Mesh.gemm_variants() for c in choice["y_pass"] + choice["z_pass"])
        assert max_rel(m.transform2d(x, 1, 0), ref) < 1e-13
        m.transform_method = method


@pytest.mark.parametrize("nz,ny", [(512, 512), (256, 384)])
def test_every_fused_tile_shape(mesh_factory, oracle_mod, nz, ny):
    """Every tile shape of the fused factored kernels that fits the grid, for
    both passes and every kind pair, against the oracle (1e-13); shapes that
    do not fit are refused."""
    from tertia_amd._lib import QssError
    m = mesh_factory(nz, ny)
    m.transform_method = 1                  # both sizes multiples of 128: the fused form
    base = m.gemm_choice()
    x = np.random.default_rng(nz + 7 * ny).standard_normal((nz, ny))
    refs = {(kz, ky): oracle_mod.transform2d(x, kz, ky, use_fft=True, nthreads=4)
            for kz in (0, 1) for ky in (0, 1)}
    tried = 0
    for v in range(64):   # indices past the last shape are refused like misfits
        for key in ("fused_y_pass", "fused_z_pass"):
            c = dict(base)
            c[key] = [v] * 8
            try:
                m.set_gemm_choice(c)
            except QssError:
                continue
            assert m.gemm_choice()[key] == [v] * 8
            tried += 1
            for (kz, ky), ref in refs.items():
                assert max_rel(m.transform2d(x, kz, ky), ref) < 1e-13, (key, v, kz, ky)
    assert tried >= 36   # 45 shapes fit 512 x 512, 39 fit 256 x 384
    m.set_gemm_choice(base)


@pytest.mark.parametrize("nz,ny", [(256, 264), (320, 256), (264, 512), (256, 256), (384, 256),
                                   (512, 384)])
def test_factored_transform_vs_oracle(mesh_factory, oracle_mod, nz, ny):
    """The factored form (one radix-2 step of the DCT-IV's half-length DFT, two
    half-size GEMMs: with the twiddles inside the GEMMs where both sizes are
    multiples of 128, between twiddle kernels otherwise) and the dense form
    both match the oracle's O(N^2)-per-row sums to 1e-13 on every kind, and
    each other."""
    from tertia_amd.params import QSS_KIND_C
    m = mesh_factory(nz, ny)
    # contexts start factored where the twiddles run inside the GEMMs (both
    # sizes multiples of 128), dense elsewhere
    assert m.transform_method == (1 if nz % 128 == 0 and ny % 128 == 0 else 0)
    m.transform_method = 1                  # FP64, multiples of 8, >= 256: allowed
    rng = np.random.default_rng(nz + ny)
    x = rng.standard_normal((nz, ny))
    for kz in (0, 1):
        for ky in (0, 1):
            ref = oracle_mod.transform2d(x, kz, ky, use_fft=True, nthreads=4)
            m.transform_method = 1
            fact = m.transform2d(x, kz, ky)
            m.transform_method = 0
            dense = m.transform2d(x, kz, ky)
            e = (max_rel(fact, ref), max_rel(dense, ref), max_rel(fact, dense))
            print("factored %dx%d kinds (%d,%d): vs oracle %.1e, dense %.1e, fact-dense %.1e"
                  % (nz, ny, kz, ky, *e), flush=True)
            assert max(e) < 1e-13, (kz, ky, e)
    m.transform_method = 0


def test_factored_method_refused_where_not_eligible(mesh_factory):
    from tertia_amd._lib import QssError
    m = mesh_factory(40, 72)
    assert m.transform_method == 0
    with pytest.raises(QssError):
        m.transform_method = 1


def test_factored_sweep_matches_dense_sweep():
    """A 256^2 sweep with either transform method: the fields agree to
    rounding (rel-L2 <= 1e-11 over 40 layers of the blowout workload)."""
    from tertia_amd.mesh import Mesh
    from tertia_amd.params import blowout_256
    p = blowout_256(nslices=40)
    out = {}
    for method in (0, 1):
        m = Mesh(p, history=("Ex", "Ey", "Bx", "Psi"))
        try:
            m.transform_method = method
            m.MoveAllSplitLayers()
            out[method] = {f: np.stack([m.history(f, s) for s in range(p.nslices)])
                           for f in ("Ex", "Ey", "Bx", "Psi")}
        finally:
            m.close()
    for f in out[0]:
        e = float(np.linalg.norm(out[1][f] - out[0][f]) / np.linalg.norm(out[0][f]))
        print("sweep factored vs dense", f, "%.1e" % e, flush=True)
        assert e <= 1e-11, (f, e)
