"""GPU parity of the evolving driver (qss_driver_seed / qss_driver_push /
qss_time_step) against the oracle on identical state.

Tolerances:
  * seeding: bitwise (same host arithmetic on both sides);
  * push: bitwise particle state (the kernel is compiled without FMA
    contraction and evaluates the reference's operation order, like the
    oracle);
  * deposit: the GPU sums in 64-bit fixed point with resolution 2^-e, e set
    from the total |w*polarity| W so that W*2^e < 2^61; each contribution is
    rounded once, so |gpu - oracle| <= (contributions per node) * 2^-(e+1)
    plus the oracle's own FP64 rounding;
  * time steps (push + sweep, several steps): max-rel 1e-9 on the Ex history
    and on particle momenta (fixed-point deposit rounding ~W*2^-61 feeds the
    sweep and then the next push).
"""
import math

import numpy as np
import pytest

from tertia_amd.params import Bunch, Driver, Params

pytestmark = pytest.mark.gpu

FIELDS6 = ("Ex", "Ey", "Ez", "Bx", "By", "Bz")


def grid(nslices=24, ny=32, nz=32, bunches=None, **kw):
    if bunches is None:
        bunches = [Bunch(nb=2.0, sigma_r=0.4, sigma_xi=0.25, xi0=1.2, polarity=-1.0, noise=0.03),
                   Bunch(nb=3.0, sigma_r=0.25, sigma_xi=0.1, xi0=1.9, y0=0.15, polarity=-1.0)]
    return Params(ny=ny, nz=nz, ly=6.0, lz=6.0, hx=0.1, nslices=nslices, n_iter=2,
                  jitter=1e-3, bunches=bunches, **kw)


def pair(oracle_mod, p, hist=FIELDS6):
    from tertia_amd.mesh import Mesh
    m = Mesh(p, history=hist)
    o = oracle_mod.Oracle(p, history=hist)
    return m, o


def resolution(weights, bunch, p):
    W = sum(abs(w * p.bunches[b].polarity) for w, b in zip(weights, bunch))
    return 2.0 ** (math.frexp(W)[1] - 61)


def random_wake(rng, p, amp=0.3):
    return {f: [amp * rng.standard_normal((p.nz, p.ny)) for _ in range(p.nslices)]
            for f in FIELDS6}


def test_seed_matches_oracle(oracle_mod):
    p = grid()
    m, o = pair(oracle_mod, p)
    try:
        drv = Driver(ts=1.0, p0=[(100.0, 0.5, -0.25), (200.0, 0, 0)], q2m=[-1.0, -1.0])
        n = m.SeedBeamParticles(drv)
        assert n == o.driver_seed(drv) and n > 0
        a, b = m.driver_particles(), o.driver_particles()
        for k in a:
            assert np.array_equal(a[k], b[k]), k
    finally:
        m.close(); o.close()


@pytest.mark.parametrize("syn", [False, True])
def test_push_matches_oracle(oracle_mod, syn):
    p = grid()
    m, o = pair(oracle_mod, p)
    rng = np.random.default_rng(5 + syn)
    try:
        drv = Driver(ts=0.7, p0=[(0, 0, 0)] * 2, q2m=[-1.0, -0.5], b_ext=(0.0, 0.01, -0.02),
                     syn_min=1e2 if syn else 1e300, wavelength=1e-4)
        m.SeedBeamParticles(drv)
        o.driver_seed(drv)
        wake = random_wake(rng, p)
        for f in FIELDS6:
            for s in range(p.nslices):
                m.set_history(f, s, wake[f][s])
                o.set_history(f, s, wake[f][s])
        n = 4000
        L = p.nslices
        cell = np.stack([rng.integers(0, L, n), rng.integers(0, p.ny, n),
                         rng.integers(0, p.nz, n)], 1)
        # ghosts and edges: clamped cells, front layer, last transverse cell
        cell[:40, 1] = -1
        cell[40:80, 0] = L
        cell[80:200, 1] = p.ny - 1
        cell[200:300, 0] = 0
        frac = rng.uniform(0, 1, (n, 3))
        frac[300:320] = 0.0
        frac[320:340] = 1.0
        mom = np.stack([rng.uniform(-5, 400, n), rng.normal(0, 20, n), rng.normal(0, 20, n)], 1)
        mom[340:400, 1] = rng.choice([-300.0, 300.0], 60)   # cross several cells / boundary
        weight = rng.uniform(0.01, 1.0, n)
        bunch = rng.integers(0, 2, n).astype(np.int32)
        m.driver_set(cell, frac, mom, weight, bunch)
        o.driver_set(cell, frac, mom, weight, bunch)
        m.MoveBeamParticles()
        o.driver_push()
        a, b = m.driver_particles(), o.driver_particles()
        for k in a:
            assert np.array_equal(a[k], b[k]), k
        res = resolution(weight, bunch, p)
        for layer in range(L + 1):
            gr, gj = m.driver_density(layer)
            orr, oj = o.driver_density(layer)
            # at most 8 nodes per particle; a node sees few contributions here
            tol = 64 * res + 1e-14 * max(np.abs(orr).max(), 1e-300)
            assert np.abs(gr - orr).max() <= tol, layer
            assert np.abs(gj - oj).max() <= tol, layer
    finally:
        m.close(); o.close()


def test_zero_step_deposit_is_the_frozen_lattice(oracle_mod):
    px0 = 20.0
    vx = px0 * (1.0 / math.sqrt(1.0 + px0 * px0))
    p = grid(bunches=[Bunch(nb=2.0, sigma_r=0.4, sigma_xi=0.2, xi0=1.2, polarity=-1.0, vx=vx,
                            noise=0.05)])
    m, o = pair(oracle_mod, p)
    try:
        m.SeedBeamParticles(Driver(ts=0.0, p0=[(px0, 0, 0)], q2m=[-1.0]))
        m.MoveBeamParticles()
        d = m.driver_particles()
        res = resolution(d["weight"], d["bunch"], p)
        for layer in range(p.nslices + 1):
            gr, gj = m.driver_density(layer)
            fr, fj = o.beam_plane(layer)
            tol = 32 * res + 1e-14 * max(np.abs(fr).max(), 1e-300)
            assert np.abs(gr - fr).max() <= tol, layer
            assert np.abs(gj - fj).max() <= tol, layer
    finally:
        m.close(); o.close()


@pytest.mark.parametrize("nx_split", [1, 2])
def test_time_steps_match_oracle(oracle_mod, nx_split):
    p = grid(nx_split=nx_split)
    hist = FIELDS6 + ("Psi",)
    m, o = pair(oracle_mod, p, hist)
    try:
        drv = Driver(ts=2.0, p0=[(30.0, 0.0, 0.0), (60.0, 0.0, 0.0)], q2m=[-1.0, -1.0],
                     syn_min=20.0, wavelength=1e-3)
        m.SeedBeamParticles(drv)
        o.driver_seed(drv)
        for step in range(3):
            rg = m.Step()
            ro = o.time_step()
            assert np.all(np.isfinite(rg))
            ex_g = np.stack([m.history("Ex", s) for s in range(p.nslices)])
            ex_o = np.stack([o.history("Ex", s) for s in range(p.nslices)])
            err = np.linalg.norm(ex_g - ex_o) / np.linalg.norm(ex_o)
            assert err < 1e-9, (step, err)
            a, b = m.driver_particles(), o.driver_particles()
            assert np.array_equal(a["cell"], b["cell"]), step
            assert np.abs(a["mom"] - b["mom"]).max() <= 1e-9 * np.abs(b["mom"]).max(), step
            assert np.abs(a["frac"] - b["frac"]).max() <= 1e-9, step
        # the driver moved and lost energy in its own wake
        assert not np.array_equal(a["mom"], np.array(
            [drv.p0[bb] for bb in a["bunch"]], dtype=float))
    finally:
        m.close(); o.close()


def test_driver_errors(oracle_mod):
    from tertia_amd._lib import QssError
    p = grid()
    m, o = pair(oracle_mod, p)
    try:
        m.SeedBeamParticles(Driver(ts=1e300, p0=[(0, 0, 0)] * 2, q2m=[-1.0, -1.0]))
        one = dict(cell=[[3, 3, 3]], mom=[[0.0, 0.0, 0.0]], weight=[1.0], bunch=[0])
        m.driver_set(frac=[[0.5, 1.5, 0.5]], **one)        # fraction outside [0,1]
        with pytest.raises(QssError):
            m.MoveBeamParticles()
        m.driver_set(frac=[[0.5, 0.5, 0.5]], **one)        # (Vx-1) ts/hx unbounded
        with pytest.raises(QssError):
            m.MoveBeamParticles()
        with pytest.raises(QssError):
            m.driver_set(cell=[[3, 3, 3]], frac=[[0.5] * 3], mom=[[0.0] * 3], weight=[1.0],
                         bunch=[5])
        with pytest.raises(QssError):
            m.driver_density(p.nslices + 1)
    finally:
        m.close(); o.close()


def test_driver_sweep_in_graph_mode_is_bitwise_eager(oracle_mod):
    p = grid(ny=32, nz=32, nslices=20)
    from tertia_amd.mesh import Mesh
    outs = []
    for graph in (False, True):
        m = Mesh(p, history=FIELDS6)
        try:
            m.SeedBeamParticles(Driver(ts=1.5, p0=[(40.0, 0, 0)] * 2, q2m=[-1.0, -1.0]))
            m.enable_graph(graph)
            for _ in range(2):
                m.Step()
            outs.append(np.stack([m.history("Ex", s) for s in range(p.nslices)]))
        finally:
            m.close()
    assert np.array_equal(outs[0], outs[1])


def test_time_steps_reference_shape_match_oracle(oracle_mod):
    """The same time steps with the reference's linked NGP plasma shape (the
    input-deck default)."""
    from tertia_amd.params import QSS_SHAPE_NGP_REF
    p = grid(shape=QSS_SHAPE_NGP_REF)
    m, o = pair(oracle_mod, p)
    try:
        drv = Driver(ts=2.0, p0=[(30.0, 0.0, 0.0), (60.0, 0.0, 0.0)], q2m=[-1.0, -1.0])
        m.SeedBeamParticles(drv)
        o.driver_seed(drv)
        for step in range(2):
            m.Step()
            o.time_step()
            ex_g = np.stack([m.history("Ex", s) for s in range(p.nslices)])
            ex_o = np.stack([o.history("Ex", s) for s in range(p.nslices)])
            assert np.linalg.norm(ex_g - ex_o) / np.linalg.norm(ex_o) < 1e-9, step
    finally:
        m.close(); o.close()


def test_driver_in_split_bf16_is_close_to_fp64():
    """bf16x6 transforms with the driver's deposit as the beam source: the
    producer of the beam planes writes their split operands too."""
    from tertia_amd.mesh import Mesh
    from tertia_amd.params import QSS_PREC_MIXED_BF16X6
    outs = []
    for prec in (0, QSS_PREC_MIXED_BF16X6):
        m = Mesh(grid(precision=prec), history=FIELDS6)
        try:
            m.SeedBeamParticles(Driver(ts=2.0, p0=[(30.0, 0, 0), (60.0, 0, 0)],
                                       q2m=[-1.0, -1.0]))
            for _ in range(2):
                m.Step()
            outs.append(np.stack([m.history("Ex", s) for s in range(m.params.nslices)]))
        finally:
            m.close()
    err = np.linalg.norm(outs[1] - outs[0]) / np.linalg.norm(outs[0])
    assert err < 1e-5, err
