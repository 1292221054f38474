"""Shared helpers of the parity tests."""
import numpy as np


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    if den == 0.0:
        return float(np.linalg.norm(a))
    return float(np.linalg.norm(a - b) / den)


def max_rel(a, b):
    """max |a-b| / max |b| (the reference's residual style)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = float(np.abs(b).max())
    if den == 0.0:
        return float(np.abs(a).max())
    return float(np.abs(a - b).max() / den)


def smooth_field(rng, nz, ny, amp=1.0, modes=4):
    """A random smooth plane: a few low Fourier modes."""
    k = np.arange(nz)[:, None] / nz
    j = np.arange(ny)[None, :] / ny
    f = np.zeros((nz, ny))
    for _ in range(modes):
        a, b = rng.integers(1, 4, size=2)
        ph = rng.uniform(0, 2 * np.pi, size=2)
        f += rng.uniform(-1, 1) * np.cos(2 * np.pi * a * j + ph[0]) * np.cos(2 * np.pi * b * k + ph[1])
    return amp * f / modes


def random_particles(rng, ny, nz, n, pmax=0.5, hot_fraction=0.0, hot_p=6.0):
    cj = rng.integers(0, ny, size=n).astype(np.int32)
    ck = rng.integers(0, nz, size=n).astype(np.int32)
    order = np.lexsort((np.arange(n), cj + ny * ck))
    cj, ck = cj[order], ck[order]
    fy = rng.uniform(0, 1, size=n)
    fz = rng.uniform(0, 1, size=n)
    p = rng.uniform(-pmax, pmax, size=(3, n))
    if hot_fraction > 0:
        hot = rng.uniform(size=n) < hot_fraction
        p[1, hot] = rng.choice([-1, 1], size=hot.sum()) * hot_p
        p[2, hot] = rng.uniform(-hot_p, hot_p, size=hot.sum())
    pid = np.arange(n, dtype=np.int64)
    return cj, ck, fy, fz, p[0], p[1], p[2], pid
