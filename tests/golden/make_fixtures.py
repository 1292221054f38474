"""Writes tests/golden/transforms.npz: half-integer transform fixtures.

Inputs are the reference's own known-answer recipes (commented-out test
mains that store no outputs):
  * m[i][j] = sin(100*i + j + 1)          CUDA_WRAP/1d_batch.cu:536-541
  * 10*i + j and -10*i + j, 2-D r2r with (RODFT11, REDFT11) and
    (REDFT11, RODFT11) plans              CUDA_WRAP/surfaceFFT.cu:400-413
Outputs come from scipy.fft dct/dst(type=4, norm=None), which is FFTW's
REDFT11/RODFT11 definition Y_k = 2 sum_j X_j cos|sin(pi (j+1/2)(k+1/2)/N)
(a library routine, not the reference).  2-D transforms are separable:
kind_z along axis 0 (rows k), kind_y along axis 1 (j fastest), exactly
fftw_plan_r2r_2d(nz, ny, in, out, kind_z, kind_y).

Run: python tests/golden/make_fixtures.py
"""
import os

import numpy as np
from scipy.fft import dct, dst

HERE = os.path.dirname(os.path.abspath(__file__))


def t1(x, kind, axis):
    return (dct if kind == 0 else dst)(x, type=4, axis=axis)


def t2(x, kz, ky):
    return t1(t1(x, ky, 1), kz, 0)


def main():
    out = {}
    for nz, ny in ((8, 16), (16, 8), (32, 32), (12, 20)):
        i = np.arange(nz)[:, None]
        j = np.arange(ny)[None, :]
        x = np.sin(100.0 * i + j + 1.0)
        out["sin_%dx%d_in" % (nz, ny)] = x
        for kz in (0, 1):
            for ky in (0, 1):
                out["sin_%dx%d_k%d%d" % (nz, ny, kz, ky)] = t2(x, kz, ky)
    n = 16
    i = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    a = 10.0 * i + j
    b = -10.0 * i + j
    out["lin_a_in"] = a
    out["lin_b_in"] = b
    out["lin_a_k10"] = t2(a, 1, 0)
    out["lin_b_k01"] = t2(b, 0, 1)
    for n in (1, 2, 3, 7, 8, 64, 96):
        x = np.sin(100.0 * np.arange(n) + 1.0)
        out["v%d_in" % n] = x
        out["v%d_c" % n] = t1(x, 0, 0)
        out["v%d_s" % n] = t1(x, 1, 0)
    np.savez_compressed(os.path.join(HERE, "transforms.npz"), **out)


if __name__ == "__main__":
    main()
