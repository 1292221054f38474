"""GPU: a reference-style deck (tests/test_deck.py's DECK: 64x64 cells,
200 layers, an electron driver of nb = 0.5, the reference's NGP split path)
swept end to end through the C ABI and compared with the FP64 oracle run
from the same deck.  Tolerances as for the 64x64 linear wake: rel-L2 1e-9."""
import numpy as np
import pytest

from tests.test_deck import DECK
from tertia_amd.deck import load_deck

pytestmark = pytest.mark.gpu

FIELDS = ("Ex", "Ez", "By", "Psi")


@pytest.mark.parametrize("bext", ["", " BxExternal = 0.3\n ByExternal = -0.2\n BzExternal = 0.15\n"])
def test_deck_sweep_vs_oracle(oracle_mod, bext):
    """Without and with the deck's uniform external B on the plasma."""
    from tertia_amd.mesh import Mesh
    p = load_deck(DECK.replace(" Nspecies = 2\n", " Nspecies = 2\n" + bext))
    assert any(p.b_ext) == bool(bext)
    m = Mesh(p, history=FIELDS)
    o = oracle_mod.Oracle(p, use_fft=True, nthreads=8, history=FIELDS)
    try:
        rg = m.MoveAllSplitLayers()
        st, ro = o.sweep()
        assert st == 0
        assert np.abs(rg - ro).max() < 1e-7
        for f in FIELDS:
            G = np.stack([m.history(f, s) for s in range(p.nslices)])
            R = np.stack([o.history(f, s) for s in range(p.nslices)])
            e = float(np.linalg.norm(G - R) / np.linalg.norm(R))
            print("deck", f, "%.1e" % e, flush=True)
            assert e < 1e-9, (f, e)
    finally:
        m.close()
        o.close()
