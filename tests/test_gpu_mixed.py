"""GPU: the mixed-precision modes (FP32 push; FP32-MFMA or split-BF16-MFMA
transforms; FP64 particle state, deposit sums, spectral state and residual).

Transforms are compared with the FP64 oracle; sweeps with the FP64 GPU sweep
on identical seeded inputs (the FP64 GPU sweep itself matches the oracle to
1e-9, see test_gpu_sweep.py).  Tolerances, per mode:
  transform max|d|/max|ref|: mixed32 3e-6, bf16x6 3e-6, bf16x3 1e-4
  sweep rel-L2 (group-relative, every layer) of E, B, psi: 1e-3 / 1e-3 / 1e-2
The sweep numbers are printed (pytest -s) for the record."""
import numpy as np
import pytest

from tests.helpers import max_rel
from tertia_amd.params import PRECISIONS, linear_wake_64, small

pytestmark = pytest.mark.gpu

MIXED = ["mixed32", "bf16x3", "bf16x6"]
TOL_T = {"mixed32": 3e-6, "bf16x3": 1e-4, "bf16x6": 3e-6}
TOL_SWEEP = {"mixed32": 1e-3, "bf16x3": 1e-2, "bf16x6": 1e-3}


@pytest.fixture(scope="module")
def meshes():
    from tertia_amd.mesh import Mesh
    made = {}

    def get(prec, nz, ny):
        key = (prec, nz, ny)
        if key not in made:
            made[key] = Mesh(small(ny=ny, nz=nz, nslices=1, precision=PRECISIONS[prec]),
                             history=())
        return made[key]
    yield get
    for m in made.values():
        m.close()


@pytest.mark.parametrize("prec", MIXED)
@pytest.mark.parametrize("nz,ny", [(64, 64), (48, 80), (512, 512)])
def test_mixed_transform_vs_oracle(meshes, oracle_mod, prec, nz, ny):
    rng = np.random.default_rng(nz * 7 + ny)
    m = meshes(prec, nz, ny)
    x = rng.standard_normal((nz, ny))
    for kz in (0, 1):
        for ky in (0, 1):
            got = m.transform2d(x, kz, ky)
            ref = oracle_mod.transform2d(x, kz, ky, use_fft=False, nthreads=4)
            e = max_rel(got, ref)
            assert e < TOL_T[prec], (prec, kz, ky, e)


@pytest.mark.parametrize("prec", MIXED)
def test_mixed_every_tile_shape(meshes, oracle_mod, prec):
    from tertia_amd.mesh import Mesh
    nv = Mesh.gemm_variants(PRECISIONS[prec])
    assert nv > 0
    for nz, ny in ((48, 80), (512, 512)):
        m = meshes(prec, nz, ny)
        x = np.random.default_rng(3).standard_normal((nz, ny))
        ref = oracle_mod.transform2d(x, 0, 1, use_fft=False, nthreads=4)
        for v in range(nv):
            m.set_gemm_variant(v)
            e = max_rel(m.transform2d(x, 0, 1), ref)
            assert e < TOL_T[prec], (prec, nz, ny, v, e)
        m.set_gemm_variant(-1)
        choice = m.autotune(1)
        assert all(0 <= c < nv for c in choice["y_pass"] + choice["z_pass"])


def sweep_history(p, fields):
    from tertia_amd.mesh import Mesh
    m = Mesh(p, history=fields)
    try:
        res = m.MoveAllSplitLayers()
        H = {f: np.stack([m.history(f, s) for s in range(p.nslices)]) for f in fields}
        part = m.particles()
    finally:
        m.close()
    return H, part, res


GROUP = {"Ex": "E", "Ey": "E", "Ez": "E", "Bx": "B", "By": "B", "Bz": "B", "Psi": "Psi"}


def small_bext(**kw):
    """The small case in a uniform external B (the float push's B_ext term)."""
    return small(b_ext=(0.5, -0.3, 0.8), **kw)


@pytest.mark.parametrize("cfg", ["small", "linear_wake_64", "small_bext"])
def test_mixed_sweep_vs_fp64(cfg):
    fields = list(GROUP)
    make = {"small": small, "linear_wake_64": linear_wake_64, "small_bext": small_bext}[cfg]
    p64 = make()
    H64, P64, r64 = sweep_history(p64, fields)
    norm = {}
    for f in fields:
        norm[GROUP[f]] = norm.get(GROUP[f], 0.0) + float(np.sum(H64[f] ** 2))
    o64 = np.argsort(P64["id"])
    for prec in MIXED:
        H, P, r = sweep_history(make(precision=PRECISIONS[prec]), fields)
        errs = {f: float(np.linalg.norm(H[f] - H64[f]) / np.sqrt(norm[GROUP[f]]))
                for f in fields}
        om = np.argsort(P["id"])
        dp = max(float(np.abs(P[k][om] - P64[k][o64]).max()) for k in ("px", "py", "pz"))
        print(cfg, prec, {f: "%.1e" % e for f, e in errs.items()}, "max|dp| %.1e" % dp,
              "resid %.1e/%.1e" % (max(r), max(r64)))
        for f, e in errs.items():
            assert e < TOL_SWEEP[prec], (cfg, prec, f, e)
        assert np.all(np.isfinite(r))


def test_mixed_needs_tiled_grid():
    from tertia_amd._lib import QssError
    from tertia_amd.mesh import Mesh
    with pytest.raises(QssError):
        Mesh(small(ny=24, nz=32, precision=PRECISIONS["mixed32"]), history=())


@pytest.mark.parametrize("prec", ["bf16x3", "bf16x6"])
def test_split_operands_from_producers(prec):
    """In split-bf16 modes the kernels producing transform inputs (window
    combine, Guess, Iterate, beam) also write the split operands and the
    transform skips its split pass.  The same layer steps must give the same
    bits as when every transform splits its own inputs, including after the
    host overwrites a produced plane (the record of that split is dropped)."""
    from tertia_amd.mesh import Mesh
    p = small(ny=32, nz=32, nslices=6, precision=PRECISIONS[prec])
    rng = np.random.default_rng(11)
    jx = rng.standard_normal((32, 32)) * 1e-2
    outs = []
    for forget in (False, True):
        m = Mesh(p, history=())
        try:
            m.reset()
            m.InterpolateBeamLayers(p.nslices - 1, 0)
            if forget:
                m.layer_ptr("Jx")          # hands out a writable plane: records dropped
            m.GuessFieldsHydroLinLayerSplit(p.nslices - 1)
            m.MoveParticlesLayerSplit(p.nslices - 1, 0, 0)
            if forget:
                m.layer_ptr("Jx")
            m.IterateFieldsHydroLinLayerSplit(p.nslices - 1, 0, 0)
            m.MoveParticlesLayerSplit(p.nslices - 1, 0, 0)
            m.set_field("Jx", jx)              # host edit after the deposit
            m.IterateFieldsHydroLinLayerSplit(p.nslices - 1, 0, 1)
            outs.append({f: m.get_field(f) for f in ("Ex", "Ey", "Bz", "Jx")})
            # the corrector transformed the edited plane, not the deposit
            ref = m.transform2d(jx, 0, 0)
            got = m.get_spectral("Jx")
            assert np.abs(got - ref).max() <= 1e-6 * np.abs(ref).max()
        finally:
            m.close()
    for f in outs[0]:
        assert np.array_equal(outs[0][f], outs[1][f]), f
    assert np.array_equal(outs[0]["Jx"], jx)
