"""Spectral calculus on the half-integer cosine/sine basis, built on scipy's
type-4 DCT/DST (a library routine, independent of the oracle's transforms).

A plane f of shape (nz, ny), node (j, k) at y = (j + 1/2) hy, z = (k + 1/2) hz,
is expanded along each axis in cos(a_m y) (kind 0) or sin(a_m y) (kind 1) with
a_m = pi (m + 1/2) / L, m = 0..n-1.  Every operator below acts on the
coefficients exactly, so an identity of the continuous fields (a divergence,
a curl, a Laplacian) holds for the planes to rounding:

* d/dy of cos(a y) is -a sin(a y) and of sin(a y) is +a cos(a y), so a
  derivative flips the kind of its axis;
* the Laplacian multiplies every mode by -(a_y^2 + a_z^2).
"""
import numpy as np
import scipy.fft as sf


def _fwd(f, kind, axis):
    n = f.shape[axis]
    return (sf.dct if kind == 0 else sf.dst)(f, type=4, axis=axis) / n


def _inv(a, kind, axis):
    return (sf.dct if kind == 0 else sf.dst)(a, type=4, axis=axis) / 2.0


def coeffs(f, kinds):
    """Mode amplitudes of f for the kinds (kind_z, kind_y)."""
    kz, ky = kinds
    return _fwd(_fwd(np.asarray(f, dtype=np.float64), ky, 1), kz, 0)


def synth(a, kinds):
    kz, ky = kinds
    return _inv(_inv(a, ky, 1), kz, 0)


class Grid:
    def __init__(self, ny, nz, ly, lz):
        self.ny, self.nz, self.ly, self.lz = ny, nz, ly, lz
        self.ay = np.pi * (np.arange(ny) + 0.5) / ly    # along y (axis 1)
        self.az = np.pi * (np.arange(nz) + 0.5) / lz    # along z (axis 0)

    def dy(self, f, kinds):
        """(df/dy, kinds of the result)."""
        kz, ky = kinds
        a = coeffs(f, kinds)
        sign = -1.0 if ky == 0 else 1.0
        out = (kz, 1 - ky)
        return synth(sign * self.ay[None, :] * a, out), out

    def dz(self, f, kinds):
        kz, ky = kinds
        a = coeffs(f, kinds)
        sign = -1.0 if kz == 0 else 1.0
        out = (1 - kz, ky)
        return synth(sign * self.az[:, None] * a, out), out

    def lap(self, f, kinds):
        a = coeffs(f, kinds)
        k2 = self.ay[None, :] ** 2 + self.az[:, None] ** 2
        return synth(-k2 * a, kinds)

    def div(self, fy, fz, ky_kinds=(0, 1), kz_kinds=(1, 0)):
        """d fy/dy + d fz/dz for fy of the Ey kind and fz of the Ez kind."""
        a, ka = self.dy(fy, ky_kinds)
        b, kb = self.dz(fz, kz_kinds)
        assert ka == kb
        return a + b, ka

    def curl_x(self, fy, fz, ky_kinds=(0, 1), kz_kinds=(1, 0)):
        """d fz/dy - d fy/dz (x component of the transverse curl)."""
        a, ka = self.dy(fz, kz_kinds)
        b, kb = self.dz(fy, ky_kinds)
        assert ka == kb
        return a - b, ka
