"""Pins the oracle's evolving-driver restatement (Mesh::SeedBeamParticles,
Mesh::MoveBeamParticles, Mesh::MoveBeamInCell; mesh.cpp:43-64,108-276,
movebeam.cpp:14-770) with closed forms and invariants, not by retyping it:

* ts = 0 in a zero wake leaves the lattice in place, so the 3-D deposit of
  every layer equals the frozen driver's plane (a separate restatement);
* zero fields: straight-line motion of the global position (i+x, j+y, k+z)
  by ((Vx-1) ts/hx, Vy ts/hy, Vz ts/hz) across several cells;
* pure E: the reference's unscaled kick twice, dp = 2E;
* pure B: the Boris rotation keeps |p| and p.B and turns the perpendicular
  part by 2 atan(|b|), b = B q2m ts / (2 gamma);
* a wake linear in (i, j, k) is gathered exactly (distinct slopes per axis
  catch a swapped or transposed corner);
* deposit: partition of unity per particle, clamped ghost cells keep only
  their interior nodes, particles outside the interior are not pushed;
* radiation reaction: no loss for dp parallel to p, loss along -p that
  scales with dp_perp^2 and (1+p^2) and as 1/ts (its absolute constant is the
  reference's literal 4/9/137 * 1.2398e-4/Wavelength * 3/(4 PI) / 5.11e5 and
  is not pinned otherwise).
"""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle
from tertia_amd.params import Bunch, Driver, Params

FIELDS6 = ("Ex", "Ey", "Ez", "Bx", "By", "Bz")


def grid(nslices=12, ny=16, nz=16, bunches=None, **kw):
    if bunches is None:
        bunches = [Bunch(nb=1.0, sigma_r=0.5, sigma_xi=0.25, xi0=0.6, polarity=-1.0)]
    return Params(ny=ny, nz=nz, ly=4.0, lz=4.0, hx=0.1, nslices=nslices, n_iter=1,
                  jitter=0.0, bunches=bunches, **kw)


def oracle(p):
    return Oracle(p, history=FIELDS6 + ("RhoBeam", "JxBeam"))


def set_wake(o, p, fn):
    """History planes from fn(name, i, J, K) with layer i = nslices-1-slice."""
    K, J = np.meshgrid(np.arange(p.nz), np.arange(p.ny), indexing="ij")
    for s in range(p.nslices):
        i = p.nslices - 1 - s
        for f in FIELDS6:
            o.set_history(f, s, fn(f, i, J, K) * np.ones((p.nz, p.ny)))


def one(o, cell, frac, mom, w=1.0, b=0):
    o.driver_set(np.array([cell]), np.array([frac]), np.array([mom]), np.array([w]),
                 np.array([b]))


def test_zero_step_deposit_is_the_frozen_lattice():
    px0 = 10.0
    vx = px0 * (1.0 / math.sqrt(1.0 + px0 * px0))
    # both bunches end before the front cell i = nslices, which the frozen
    # lattice seeds and the reference's seeding range (i < l_Mx) does not
    bunches = [Bunch(nb=1.5, sigma_r=0.4, sigma_xi=0.2, xi0=1.3, polarity=-1.0, vx=vx,
                     noise=0.05),
               Bunch(nb=0.7, sigma_r=0.3, sigma_xi=0.15, xi0=1.6, y0=0.2, polarity=1.0,
                     vx=vx)]
    p = grid(nslices=24, bunches=bunches)
    o = oracle(p)
    n = o.driver_seed(Driver(ts=0.0, p0=[(px0, 0, 0)] * 2, q2m=[-1.0, -1.0]))
    assert n > 0 and n % 4 == 0
    before = o.driver_particles()
    o.driver_push()
    after = o.driver_particles()
    for k in ("cell", "frac", "mom"):
        assert np.array_equal(before[k], after[k]), k
    for layer in range(p.nslices + 1):
        rho, jx = o.driver_density(layer)
        frho, fjx = o.beam_plane(layer)
        scale = max(np.abs(frho).max(), 1e-300)
        assert np.abs(rho - frho).max() <= 1e-14 * scale, layer
        assert np.abs(jx - fjx).max() <= 1e-14 * scale, layer
    o.close()


@pytest.mark.parametrize("mom", [(3.0, 5.0, -4.0), (0.5, -30.0, 2.0), (-2.0, 0.1, 45.0)])
def test_straight_line_in_zero_wake(mom):
    p = grid()
    o = oracle(p)
    ts = 0.35
    o.driver_seed(Driver(ts=ts, p0=[(0, 0, 0)], q2m=[-1.0], syn_min=1e300))
    start = ((6, 7, 8), (0.25, 0.5, 0.75))
    one(o, *start, mom)
    o.driver_push()
    d = o.driver_particles()
    g = 1.0 / math.sqrt(1.0 + sum(m * m for m in mom))
    hy = p.ly / p.ny
    want = (6.25 + (mom[0] * g - 1.0) * ts / p.hx, 7.5 + mom[1] * g * ts / hy,
            8.75 + mom[2] * g * ts / hy)
    lim = (p.nslices, p.ny, p.nz)
    got = d["cell"][0] + d["frac"][0]
    for a in range(3):
        if -1 < want[a] < lim[a]:
            assert abs(got[a] - want[a]) < 1e-12, (a, got, want)
    assert np.array_equal(d["mom"][0], np.array(mom))
    o.close()


def test_pure_e_kick_is_twice_the_field():
    p = grid()
    o = oracle(p)
    E = {"Ex": 0.3, "Ey": -0.7, "Ez": 1.1}
    set_wake(o, p, lambda f, i, J, K: E.get(f, 0.0))
    o.driver_seed(Driver(ts=0.1, p0=[(0, 0, 0)], q2m=[-1.0], syn_min=1e300))
    one(o, (5, 5, 5), (0.3, 0.6, 0.2), (100.0, 1.0, -2.0))
    o.driver_push()
    m = o.driver_particles()["mom"][0]
    want = np.array([100.0 + 2 * 0.3, 1.0 - 2 * 0.7, -2.0 + 2 * 1.1])
    assert np.abs(m - want).max() < 1e-12
    o.close()


@pytest.mark.parametrize("B", [(0.0, 0.0, 2.0), (1.0, -0.5, 0.25), (-3.0, 1.0, 0.0)])
def test_pure_b_boris_rotation(B):
    p = grid()
    o = oracle(p)
    Bd = dict(zip(("Bx", "By", "Bz"), B))
    set_wake(o, p, lambda f, i, J, K: Bd.get(f, 0.0))
    ts, q2m = 0.4, -1.0
    o.driver_seed(Driver(ts=ts, p0=[(0, 0, 0)], q2m=[q2m], syn_min=1e300))
    p0 = np.array([2.0, 1.0, -0.5])   # not parallel to any B below
    one(o, (4, 9, 3), (0.5, 0.5, 0.5), tuple(p0))
    o.driver_push()
    p1 = o.driver_particles()["mom"][0]
    g = math.sqrt(1.0 + p0 @ p0)
    Bv = np.array(B)
    bhat = Bv / np.linalg.norm(Bv)
    b = np.linalg.norm(Bv) * abs(q2m) * ts / (2.0 * g)
    assert abs(np.linalg.norm(p1) - np.linalg.norm(p0)) < 1e-12
    assert abs(p1 @ bhat - p0 @ bhat) < 1e-12
    u0 = p0 - (p0 @ bhat) * bhat
    u1 = p1 - (p1 @ bhat) * bhat
    ang = math.atan2(np.linalg.norm(np.cross(u0, u1)), u0 @ u1)
    assert abs(ang - 2.0 * math.atan(b)) < 1e-11
    o.close()


def test_linear_wake_is_gathered_exactly():
    p = grid()
    o = oracle(p)
    slope = {"Ex": (0.01, 0.002, 0.0003, 0.00004), "Ey": (-0.02, 0.001, -0.004, 0.0005),
             "Ez": (0.03, -0.0007, 0.0002, 0.006)}

    def fn(f, i, J, K):
        if f not in slope:
            return 0.0
        a, b, c, d = slope[f]
        return a + b * i + c * J + d * K

    set_wake(o, p, fn)
    o.driver_seed(Driver(ts=1e-3, p0=[(0, 0, 0)], q2m=[-1.0], syn_min=1e300))
    cell, frac = (4, 6, 9), (0.37, 0.81, 0.13)
    one(o, cell, frac, (0.0, 0.0, 0.0))
    o.driver_push()
    m = o.driver_particles()["mom"][0]
    X, Y, Z = (c + f for c, f in zip(cell, frac))
    for a, f in enumerate(("Ex", "Ey", "Ez")):
        s0, si, sj, sk = slope[f]
        assert abs(m[a] - 2.0 * (s0 + si * X + sj * Y + sk * Z)) < 1e-14, f
    o.close()


def test_deposit_partition_of_unity_and_ghost_cells():
    p = grid()
    o = oracle(p)
    o.driver_seed(Driver(ts=0.2, p0=[(0, 0, 0)], q2m=[-1.0], syn_min=1e300))
    L, ny, nz = p.nslices, p.ny, p.nz
    cells = np.array([[5, 5, 5], [2, 7, 3], [6, ny - 1, 4], [7, 0, 2], [L, 3, 3], [3, -1, 4]])
    frac = np.array([[0.5, 0.5, 0.5], [0.2, 0.4, 0.6], [0.5, 0.5, 0.5], [0.5, 0.5, 0.5],
                     [0.5, 0.5, 0.5], [0.5, 0.5, 0.5]])
    # 3: crosses j = Ny (clamped to the ghost cell Ny: both nodes dropped);
    # 4: crosses j = 0 downwards (ghost cell -1: keeps node j = 0 only);
    # 5, 6: start outside the interior: neither pushed nor deposited
    mom = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 40.0, 0.0], [0.0, -40.0, 0.0],
                    [0.0, 1.0, 0.0], [0.0, 1.0, 0.0]])
    w = np.array([1.0, 2.0, 4.0, 8.0, 16.0, 32.0])
    o.driver_set(cells, frac, mom, w, np.zeros(6, dtype=np.int32))
    o.driver_push()
    d = o.driver_particles()
    assert np.array_equal(d["cell"][4:], cells[4:]) and np.array_equal(d["frac"][4:], frac[4:])
    assert d["cell"][2][1] == ny and d["cell"][3][1] == -1
    tot = sum(o.driver_density(i)[0].sum() for i in range(L + 1))
    pol = -1.0
    y3 = d["frac"][3][1]
    assert abs(tot - pol * (1.0 + 2.0 + 8.0 * y3)) < 1e-12
    o.close()


def test_radiation_reaction_direction_and_scaling():
    p = grid()

    def run(B, ex=0.0, syn=True, ts=0.1, mom=(500.0, 0.0, 0.0), lam=1e-3):
        o = oracle(p)
        set_wake(o, p, lambda f, i, J, K: {"By": B, "Ex": ex}.get(f, 0.0))
        o.driver_seed(Driver(ts=ts, p0=[(0, 0, 0)], q2m=[-1.0],
                             syn_min=1e2 if syn else 1e300, wavelength=lam))
        one(o, (5, 5, 5), (0.5, 0.5, 0.5), mom)
        o.driver_push()
        m = o.driver_particles()["mom"][0]
        o.close()
        return m

    # dp parallel to p: no perpendicular change, no loss
    assert np.array_equal(run(0.0, ex=0.5), run(0.0, ex=0.5, syn=False))
    loss = []
    # small rotation angles (b ~ 1e-4), losses ~ 1e-6 |p| (well above rounding)
    for B, ts, lam, mom in ((1.0, 0.1, 1e-5, 500.0), (2.0, 0.1, 1e-5, 500.0),
                            (1.0, 0.2, 1e-5, 500.0), (1.0, 0.1, 2e-5, 500.0),
                            (1.0, 0.1, 1e-5, 1000.0)):
        on = run(B, ts=ts, lam=lam, mom=(mom, 0, 0))
        off = run(B, syn=False, ts=ts, lam=lam, mom=(mom, 0, 0))
        # radiated along -p: same direction, smaller magnitude
        assert np.linalg.norm(np.cross(on, off)) < 1e-9 * np.linalg.norm(off) ** 2
        loss.append(np.linalg.norm(off) - np.linalg.norm(on))
    base = loss[0]
    assert base > 0
    # dp_perp = p * theta, theta = 2 atan(B ts / (2 gamma)) ~ B ts / gamma:
    # loss ~ (1+p^2) dp_perp^2 / ts
    assert abs(loss[1] / base - 4.0) < 1e-4          # B doubled
    assert abs(loss[2] / base - 2.0) < 1e-4          # ts doubled: dp^2 x4, 1/ts
    assert abs(loss[3] / base - 0.5) < 1e-6          # wavelength doubled
    # p doubled: (1+p^2) x4, dp_perp = p theta with theta ~ 1/gamma: unchanged
    assert abs(loss[4] / base - 4.0) < 1e-4


@pytest.mark.parametrize("nx_split", [1, 2])
def test_time_step_at_zero_ts_is_the_frozen_sweep(nx_split):
    """Domain::Step = beam push + sweep.  With ts = 0 in a zero wake the
    deposit is the frozen lattice, so the sweep that reads the 3-D deposit
    (with the split interpolation of layers i, i+1) equals the frozen
    driver's sweep."""
    px0 = 20.0
    vx = px0 * (1.0 / math.sqrt(1.0 + px0 * px0))
    bunches = [Bunch(nb=2.0, sigma_r=0.4, sigma_xi=0.2, xi0=1.3, polarity=-1.0, vx=vx)]
    p = grid(nslices=24, bunches=bunches, nx_split=nx_split)
    frozen = oracle(p)
    frozen.sweep()
    o = oracle(p)
    o.driver_seed(Driver(ts=0.0, p0=[(px0, 0, 0)], q2m=[-1.0]))
    o.time_step()
    for s in range(p.nslices):
        a, b = o.history("Ex", s), frozen.history("Ex", s)
        assert np.abs(a - b).max() <= 1e-12 * max(np.abs(b).max(), 1e-30), s
    frozen.close()
    o.close()
