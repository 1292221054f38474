"""Pins the oracle's half-integer transforms (FFTW REDFT11/RODFT11,
movefieldshydrolinlayersplit.cpp:389-427) without the reference binary:

* scipy type-4 DCT/DST fixtures on the reference's own recipe inputs
  (tests/golden/transforms.npz, written by tests/golden/make_fixtures.py);
* brute-force sums written independently here (numpy, direct definition);
* the involution T.T = 2N.I per axis (so the inverse is the same transform
  divided by 4 Ny Nz, normalization.cu:241);
* single-mode known answers: the basis vector of mode k0 maps to N at k0 and
  0 elsewhere (orthogonality of the half-integer cosines/sines).
A transposed operand, a wrong kind per axis or a dropped factor 2 fails at
least one of these.
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "transforms.npz")


def brute1d(x, kind):
    n = len(x)
    j = np.arange(n)[:, None]
    k = np.arange(n)[None, :]
    a = np.pi * (j + 0.5) * (k + 0.5) / n
    m = 2 * (np.cos(a) if kind == 0 else np.sin(a))
    return x @ m


@pytest.mark.parametrize("use_fft", [False, True])
def test_1d_fixtures(oracle_mod, use_fft):
    g = np.load(GOLD)
    for n in (1, 2, 3, 7, 8, 64, 96):
        x = g["v%d_in" % n]
        for kind, key in ((0, "c"), (1, "s")):
            got = oracle_mod.transform1d(x, kind, use_fft)
            ref = g["v%d_%s" % (n, key)]
            assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max()), (n, kind)


@pytest.mark.parametrize("use_fft", [False, True])
def test_2d_recipe_fixtures(oracle_mod, use_fft):
    g = np.load(GOLD)
    for nz, ny in ((8, 16), (16, 8), (32, 32), (12, 20)):
        x = g["sin_%dx%d_in" % (nz, ny)]
        for kz in (0, 1):
            for ky in (0, 1):
                ref = g["sin_%dx%d_k%d%d" % (nz, ny, kz, ky)]
                got = oracle_mod.transform2d(x, kz, ky, use_fft=use_fft)
                assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), (nz, ny, kz, ky)
    assert np.allclose(oracle_mod.transform2d(g["lin_a_in"], 1, 0), g["lin_a_k10"],
                       rtol=0, atol=1e-11 * np.abs(g["lin_a_k10"]).max())
    assert np.allclose(oracle_mod.transform2d(g["lin_b_in"], 0, 1), g["lin_b_k01"],
                       rtol=0, atol=1e-11 * np.abs(g["lin_b_k01"]).max())


def test_brute_force_and_fft_agree(oracle_mod):
    rng = np.random.default_rng(0)
    for n in (4, 6, 16, 32, 128):
        x = rng.standard_normal(n)
        for kind in (0, 1):
            b = brute1d(x, kind)
            for use_fft in (False, True):
                got = oracle_mod.transform1d(x, kind, use_fft)
                assert np.abs(got - b).max() < 1e-12 * max(1, np.abs(b).max()), (n, kind, use_fft)


def test_separable_2d_against_brute(oracle_mod):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((10, 14))
    for kz in (0, 1):
        for ky in (0, 1):
            ref = np.stack([brute1d(r, ky) for r in x])          # along y (rows)
            ref = np.stack([brute1d(c, kz) for c in ref.T]).T    # along z (columns)
            assert np.abs(oracle_mod.transform2d(x, kz, ky) - ref).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("use_fft", [False, True])
def test_involution(oracle_mod, use_fft):
    rng = np.random.default_rng(2)
    nz, ny = 32, 64
    x = rng.standard_normal((nz, ny))
    for kz in (0, 1):
        for ky in (0, 1):
            y = oracle_mod.transform2d(x, kz, ky, use_fft=use_fft)
            back = oracle_mod.transform2d(y, kz, ky, scale=1.0 / (4 * ny * nz), use_fft=use_fft)
            assert np.abs(back - x).max() < 1e-13


def test_single_mode(oracle_mod):
    n = 24
    j = np.arange(n)
    for kind in (0, 1):
        for k0 in (0, 5, n - 1):
            a = np.pi * (j + 0.5) * (k0 + 0.5) / n
            basis = np.cos(a) if kind == 0 else np.sin(a)
            y = oracle_mod.transform1d(basis, kind)
            expect = np.zeros(n)
            expect[k0] = n
            assert np.abs(y - expect).max() < 1e-12, (kind, k0)
