"""Pins the oracle's frozen driver planes (Mesh::SeedBeamParticles,
mesh.cpp:108-276; Specie::Density case 1, plasma.cpp:350-377;
Mesh::MoveBeamInCell, movebeam.cpp:662-770) against the analytic
bi-Gaussian, not against another restatement:

* charge of every layer: polarity * nb * 2 pi sigma_r^2 times the mean of the
  longitudinal profile at the two cells that deposit into the layer (cell i at
  xi = (L - i) hx and cell i - 1), with the reference's R = sqrt(2) sigma
  (exp(-d^2/R^2) = exp(-d^2 / 2 sigma^2)); the transverse sum of a well
  resolved Gaussian on a uniform grid is exact to rounding, and the 3-D cut
  at arg = 16 removes at most exp(-(16 - a_x)) of a layer whose longitudinal
  argument is a_x;
* transverse centroid y0 + hy/2, z0 + hz/2 and variance sigma_r^2 + h^2/4: the
  macro-particles sit at {1/3, 2/3} of a cell but carry the density of its
  corner, so node j collects the corners j-1 and j (a half-cell offset and a
  two-point mean that a change of convention would break);
* Jx = vx * rho node by node;
* the piecewise-constant profile (bins) replaces the longitudinal factor of
  bunch 0 by the bin value at each depositing cell.
"""
import math

import numpy as np
import pytest

from tertia_amd.params import Bunch, Params


def grid(bunches, nslices=100, **kw):
    return Params(ny=64, nz=56, ly=8.0, lz=7.0, hx=0.05, nslices=nslices, jitter=0.0,
                  bunches=bunches, **kw)


def layer_stats(o, p, layer):
    rho, jx = o.beam_plane(layer)
    hy, hz = p.ly / p.ny, p.lz / p.nz
    y = np.arange(p.ny) * hy - p.ly / 2          # node j at the corner of cell j
    z = np.arange(p.nz) * hz - p.lz / 2
    q = rho.sum()
    if q == 0.0:
        return rho, jx, 0.0, None
    my = (rho.sum(0) * y).sum() / q
    mz = (rho.sum(1) * z).sum() / q
    vy = (rho.sum(0) * (y - my) ** 2).sum() / q
    vz = (rho.sum(1) * (z - mz) ** 2).sum() / q
    return rho, jx, q * hy * hz, (my, mz, vy, vz)


def check_beam(make):
    b = Bunch(nb=1.3, sigma_r=0.45, sigma_xi=0.5, xi0=2.5, y0=0.4, z0=-0.3,
              polarity=-1.0, vx=0.9)
    p = grid([b])
    o = make(p)
    L = p.nslices
    hy, hz = p.ly / p.ny, p.lz / p.nz
    gx = lambda xi: math.exp(-(xi - b.xi0) ** 2 / (2 * b.sigma_xi ** 2))
    checked = 0
    for layer in range(1, L):
        rho, jx, q, mom = layer_stats(o, p, layer)
        want = b.polarity * b.nb * 2 * math.pi * b.sigma_r ** 2 * 0.5 * (
            gx((L - layer) * p.hx) + gx((L - layer + 1) * p.hx))
        if abs(want) < 1e-3:
            continue
        checked += 1
        ax = max((xi - b.xi0) ** 2 / (2 * b.sigma_xi ** 2)
                 for xi in ((L - layer) * p.hx, (L - layer + 1) * p.hx))
        cut = 2 * math.exp(-(16 - ax))
        assert abs(q - want) <= (1e-6 + cut) * abs(want), ("charge", layer, q, want)
        assert np.abs(jx - b.vx * rho).max() <= 1e-15 * np.abs(rho).max()
        if ax > 2.0:
            continue
        my, mz, vy, vz = mom
        assert abs(my - (b.y0 + hy / 2)) < 1e-6 and abs(mz - (b.z0 + hz / 2)) < 1e-6, mom
        # the cut at r^2 = 32 sigma^2 shifts the variance by ~ 32 exp(-14) ~ 3e-5
        assert abs(vy - (b.sigma_r ** 2 + hy ** 2 / 4)) < 1e-4 * b.sigma_r ** 2, mom
        assert abs(vz - (b.sigma_r ** 2 + hz ** 2 / 4)) < 1e-4 * b.sigma_r ** 2, mom
    o.close()
    assert checked > 20


def check_bins(make):
    nbins = 12
    bins = np.linspace(0.5, 2.0, nbins) * (1 + 0.3 * np.cos(np.arange(nbins)))
    b = Bunch(nb=3.0, sigma_r=0.5, sigma_xi=0.8, xi0=2.0)
    p = grid([b], bins=list(bins), bin_xi0=0.5, bin_xi1=3.5)
    o = make(p)
    L = p.nslices

    def binv(xi):
        t = (xi - p.bin_xi0) / (p.bin_xi1 - p.bin_xi0) * nbins
        return bins[int(t)] if 0 <= t < nbins else 0.0

    for layer in range(1, L):
        _, _, q, _ = layer_stats(o, p, layer)
        want = -1.0 * 2 * math.pi * b.sigma_r ** 2 * 0.5 * (
            binv((L - layer) * p.hx) + binv((L - layer + 1) * p.hx))
        assert abs(q - want) <= 1e-6 * max(abs(want), 1e-3), (layer, q, want)
    o.close()


def test_beam_plane_analytic(oracle_mod):
    check_beam(oracle_mod.Oracle)


def test_beam_bins_analytic(oracle_mod):
    check_bins(oracle_mod.Oracle)
