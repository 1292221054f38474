"""GPU: the adjoint xi sweep (qss_sweep_checkpointed + qss_sweep_vjp) against
central finite differences of the FP64 forward sweep.

Loss: -mean(Ex) + 1e-3 var(Ex) over a witness window of layers and radii
(Ex = the longitudinal field).  Gradient with respect to the driver's
piecewise-constant longitudinal profile.  Tolerance: |g_adj - g_fd| <=
1e-4 max|g| per checked bin (central differences, h = 1e-6 relative)."""
import os
import subprocess
import sys

import numpy as np
import pytest

from tertia_amd.loss import witness_loss_np
from tertia_amd.params import Bunch, Params, profile_gradient_256

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def small_bins(**kw):
    nb = 16
    x0, x1 = 0.0, 0.64
    w = (x1 - x0) / nb
    bins = [2.0 * np.exp(-((x0 + (i + 0.5) * w - 0.32) ** 2) / (2 * 0.15 ** 2)) + 0.1
            for i in range(nb)]
    p = Params(ny=32, nz=32, ly=4.0, lz=4.0, hx=0.02, nslices=64,
               bunches=[Bunch(nb=2.0, sigma_r=0.4, sigma_xi=0.15, xi0=0.32)],
               bins=bins, bin_xi0=x0, bin_xi1=x1)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


WIN_SMALL = dict(xi0=0.9, xi1=1.1, rmax=0.6)


def loss_of(m, p, nrun, win):
    m.reset()
    m.sweep_range(0, nrun)
    ex = np.stack([m.history("Ex", s) for s in range(nrun)])
    return witness_loss_np(ex, p, **win)


def adjoint_grad(m, p, nrun, every, win):
    m.sweep_checkpointed(nrun, every)
    ex = np.stack([m.history("Ex", s) for s in range(nrun)])
    loss, gex = witness_loss_np(ex, p, **win)
    return loss, m.sweep_vjp(gex), ex


def fd_grad(m, p, nrun, win, idx, rel=1e-6):
    base = np.array(p.bins, dtype=float)
    out = []
    for i in idx:
        h = rel * max(1.0, abs(base[i]))
        b = base.copy()
        b[i] += h
        m.set_bins(b)
        lp, _ = loss_of(m, p, nrun, win)
        b[i] -= 2 * h
        m.set_bins(b)
        lm, _ = loss_of(m, p, nrun, win)
        out.append((lp - lm) / (2 * h))
    m.set_bins(base)
    return np.array(out)


def test_checkpointed_forward_is_the_forward():
    from tertia_amd.mesh import Mesh
    p = small_bins()
    m = Mesh(p, history=("Ex",))
    try:
        m.reset()
        m.sweep_range(0, 56)
        a = np.stack([m.history("Ex", s) for s in range(56)])
        m.sweep_checkpointed(56, 8)
        b = np.stack([m.history("Ex", s) for s in range(56)])
    finally:
        m.close()
    assert np.array_equal(a, b)


@pytest.mark.parametrize("every", [8, 6])
def test_graph_replay_is_eager(every):
    """Checkpointed forward by graph pairs and the reverse by one graph per
    full checkpoint segment against the same sweeps launched eagerly:
    histories and gradients bitwise equal (every = 6: a partial segment of
    two layers first in the reverse)."""
    from tertia_amd.mesh import Mesh
    p = small_bins()
    nrun = 56
    out = []
    for graph in (True, False):
        m = Mesh(p, history=("Ex",))
        try:
            m.enable_graph(graph)
            loss, g, ex = adjoint_grad(m, p, nrun, every, WIN_SMALL)
            g2 = m.sweep_vjp(witness_loss_np(ex, p, **WIN_SMALL)[1])   # replayed again
            out.append((ex, g, g2))
        finally:
            m.close()
    (ex_g, g_g, g2_g), (ex_e, g_e, g2_e) = out
    assert np.abs(g_e).max() > 0
    assert np.array_equal(ex_g, ex_e)
    assert np.array_equal(g_g, g_e)
    assert np.array_equal(g2_g, g_e) and np.array_equal(g2_e, g_e)


def test_taped_forward_is_the_checkpointed():
    """every = 0: the forward records every layer's tape and the reverse
    recomputes nothing; histories and gradient bitwise those of the
    checkpointed sweep (every = 8), and the sweep is the forward."""
    from tertia_amd.mesh import Mesh
    p = small_bins()
    nrun = 56
    m = Mesh(p, history=("Ex",))
    try:
        m.reset()
        m.sweep_range(0, nrun)
        ex0 = np.stack([m.history("Ex", s) for s in range(nrun)])
        _, g8, ex8 = adjoint_grad(m, p, nrun, 8, WIN_SMALL)
        _, g0, ex0t = adjoint_grad(m, p, nrun, 0, WIN_SMALL)
        g0b = m.sweep_vjp(witness_loss_np(ex0t, p, **WIN_SMALL)[1])   # the tape is kept
    finally:
        m.close()
    assert np.abs(g8).max() > 0
    assert np.array_equal(ex0, ex8) and np.array_equal(ex0, ex0t)
    assert np.array_equal(g0, g8) and np.array_equal(g0b, g8)


def test_zero_loss_adjoint_is_zero():
    from tertia_amd.mesh import Mesh
    p = small_bins()
    m = Mesh(p, history=("Ex",))
    try:
        m.sweep_checkpointed(40, 8)
        g = m.sweep_vjp(None)
    finally:
        m.close()
    assert np.all(g == 0.0)


@pytest.mark.parametrize("every,b_ext", [(8, (0.0, 0.0, 0.0)), (56, (0.0, 0.0, 0.0)),
                                          (8, (0.5, -0.3, 0.8))])
def test_adjoint_vs_finite_differences_small(every, b_ext):
    """The last case puts the plasma in a uniform external B."""
    from tertia_amd.mesh import Mesh
    p = small_bins(b_ext=b_ext)
    nrun = 56
    m = Mesh(p, history=("Ex",))
    try:
        loss, g, _ = adjoint_grad(m, p, nrun, every, WIN_SMALL)
        idx = [2, 5, 8, 11, 14]
        gf = fd_grad(m, p, nrun, WIN_SMALL, idx)
    finally:
        m.close()
    scale = np.abs(g).max()
    err = np.abs(g[idx] - gf) / scale
    print("small every=%d loss %.6e |g|max %.3e adj %s fd %s err %.2e" %
          (every, loss, scale, np.array2string(g[idx], precision=6),
           np.array2string(gf, precision=6), err.max()))
    assert scale > 0
    assert err.max() < 1e-4, err


def test_adjoint_vs_finite_differences_256():
    """256x256 x 4 ppc, 650 layers (window xi in [6, 6.5], r < 0.2), 4 bins."""
    from tertia_amd.mesh import Mesh
    p = profile_gradient_256()
    nrun = 650
    win = dict(xi0=6.0, xi1=6.5, rmax=0.2)
    m = Mesh(p, history=("Ex",))
    try:
        loss, g, _ = adjoint_grad(m, p, nrun, 32, win)
        rng = np.random.default_rng(4)
        idx = sorted(rng.choice(np.arange(40, 200), 4, replace=False).tolist())
        gf = fd_grad(m, p, nrun, win, idx)
    finally:
        m.close()
    scale = np.abs(g).max()
    err = np.abs(g[idx] - gf) / scale
    print("profile_gradient_256 loss %.6e |g|max %.3e adj %s fd %s err %.2e" %
          (loss, scale, np.array2string(g[idx], precision=6),
           np.array2string(gf, precision=6), err.max()))
    assert err.max() < 1e-4, err


def test_torch_autograd_wrapper():
    """torch.autograd.Function over the adjoint, in a child process that
    imports torch before the library (one HIP runtime)."""
    code = r"""
import sys, numpy as np, torch
sys.path.insert(0, %r)
from tests.test_gpu_adjoint import small_bins, WIN_SMALL
from tertia_amd.mesh import Mesh
from tertia_amd.autograd import SweepEx, witness_loss, witness_loss_np
p = small_bins()
m = Mesh(p, history=("Ex",))
bins = torch.tensor(p.bins, dtype=torch.float64, requires_grad=True)
ex = SweepEx.apply(bins, m, 56, 8)
loss = witness_loss(ex, p, **WIN_SMALL)
loss.backward()
l2, gex = witness_loss_np(ex.detach().cpu().numpy(), p, **WIN_SMALL)
m.sweep_checkpointed(56, 8)
g2 = m.sweep_vjp(gex)
# a second forward on the same mesh before the first backward: that backward
# must refuse (it would return the second forward's gradient)
ba = torch.tensor(p.bins, dtype=torch.float64, requires_grad=True)
la = witness_loss(SweepEx.apply(ba, m, 56, 8), p, **WIN_SMALL)
lb = witness_loss(SweepEx.apply(ba * 1.01, m, 56, 8), p, **WIN_SMALL)
try:
    la.backward()
    stale_refused = False
except RuntimeError:
    stale_refused = True
m.close()
d = float(np.abs(bins.grad.numpy() - g2).max() / np.abs(g2).max())
print("torch loss %%.6e numpy loss %%.6e grad diff %%.2e stale refused %%s" %%
      (loss.item(), l2, d, stale_refused))
sys.exit(0 if abs(loss.item() - l2) <= 1e-12 * abs(l2) and d < 1e-12 and stale_refused
         else 1)
""" % ROOT
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                       timeout=300, cwd=ROOT)
    print(r.stdout, r.stderr[-2000:])
    assert r.returncode == 0
