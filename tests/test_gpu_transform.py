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
        for v in range(Mesh.gemm_variants()):
            m.set_gemm_variant(v)
            assert max_rel(m.transform2d(x, 1, 0), ref) < 1e-13, (nz, ny, v)
        m.set_gemm_variant(-1)
        choice = m.autotune(2)
        assert all(0 <= c < Mesh.gemm_variants() for c in choice["y_pass"] + choice["z_pass"])
        assert max_rel(m.transform2d(x, 1, 0), ref) < 1e-13
