"""qss_transform2d_noalloc (the context-free transform whose tables and
intermediate live in a fixed per-device scratch array) against the context's
qss_transform2d, through the C ABI, on two calls in a row, the second on a
smaller plane: for every transform kind, max |noalloc - context| <= 1e-13
max |context| (the tables of the two paths may differ in the last bit of a
cos/sin; the GEMM tile shape is the same)."""
import ctypes

import numpy as np
import pytest

from tertia_amd.params import small

pytestmark = pytest.mark.gpu


def one_size(ny, nz, seed):
    from tertia_amd.mesh import Mesh
    m = Mesh(small(ny=ny, nz=nz, nslices=2), history=())
    try:
        rng = np.random.default_rng(seed)
        for kz in (0, 1):
            for ky in (0, 1):
                x = rng.standard_normal((nz, ny))
                scale = 1.0 / (4.0 * ny * nz) if (kz + ky) % 2 else 0.75
                m.set_field("Rho", x)
                rc = m.L.qss_transform2d(m.h, kz, ky, m.layer_ptr("Rho"), m.layer_ptr("Psi"),
                                         scale)
                assert rc == 0
                ref = m.get_field("Psi")
                rc = m.L.qss_transform2d_noalloc(ny, nz, kz, ky, m.layer_ptr("Rho"),
                                                 m.layer_ptr("Jx"), scale,
                                                 ctypes.c_void_p(m.stream))
                assert rc == 0, m.L.qss_last_error()
                got = m.get_field("Jx")   # synchronises the context's stream
                err = float(np.abs(got - ref).max() / np.abs(ref).max())
                print("noalloc %dx%d kinds (%d,%d): %.1e" % (ny, nz, kz, ky, err), flush=True)
                assert err <= 1e-13, (ny, nz, kz, ky, err)
    finally:
        m.close()


def test_noalloc_transform_matches_the_context_transform():
    one_size(512, 384, 1)
    one_size(96, 80, 2)   # second call, smaller plane: the scratch keeps the first's contents


def test_noalloc_refuses_planes_beyond_the_scratch():
    from tertia_amd import _lib
    L = _lib.load()
    assert L.qss_transform2d_noalloc(513, 64, 0, 0, None, None, 1.0, None) == -1
    assert b"512" in L.qss_last_error()
