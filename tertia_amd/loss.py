"""Witness-window loss of the longitudinal wake field (numpy), used by the
adjoint checks and the benchmark; tertia_amd.autograd has the torch form."""
from __future__ import annotations

import numpy as np


def window_mask(p, nrun: int, xi0: float, xi1: float, rmax: float) -> np.ndarray:
    """Boolean (nrun, nz, ny): layers whose xi = (s+1) hx lies in [xi0, xi1]
    and cell centres ((j+1/2) hy - ly/2, (k+1/2) hz - lz/2) with r < rmax."""
    xi = (np.arange(nrun) + 1) * p.hx
    ys = (np.arange(p.ny) + 0.5) * (p.ly / p.ny) - p.ly / 2
    zs = (np.arange(p.nz) + 0.5) * (p.lz / p.nz) - p.lz / 2
    r = np.sqrt(ys[None, :] ** 2 + zs[:, None] ** 2)
    m = ((xi >= xi0 - 1e-12) & (xi <= xi1 + 1e-12))[:, None, None] & (r < rmax)[None]
    return m


def witness_loss_np(ex: np.ndarray, p, xi0: float = 6.0, xi1: float = 6.5, rmax: float = 0.2,
                    var_weight: float = 1e-3):
    """The same loss and its gradient with respect to ex, in numpy."""
    m = window_mask(p, ex.shape[0], xi0, xi1, rmax)
    v = ex[m]
    n = v.size
    mean = v.mean()
    loss = -mean + var_weight * ((v - mean) ** 2).mean()
    g = np.zeros_like(ex)
    g[m] = -1.0 / n + var_weight * 2.0 * (v - mean) / n
    return float(loss), g
