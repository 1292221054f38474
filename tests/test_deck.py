"""Reference-style input decks: the NList reading rules of namelist.cpp:93-146
and the mapping of &Domain / &Electrons / &Specie<i> to the sweep's
parameters (grid sizing of domain.cpp:141-145, the Gaussian-ellipsoid beam of
plasma.cpp:350-377, the float read of doubles)."""
import math

import numpy as np
import pytest

from tertia_amd.deck import DeckError, load_deck, read_nlist
from tertia_amd.params import QSS_SHAPE_CIC, QSS_SHAPE_NGP_REF

DECK = """# a reference-style deck
&Domain
 Hx = 0.05
 Hy = 0.1875
 Hz = 0.1875
 Xlength = 9.99
 Ylength = 11.9
 Zlength = 11.9
 NxSplit = 1
 nIter = 2
 Nspecies = 2
/
&Electrons
 Density = 1.
 P_perCell = 4
 End = 1e6
 PlateauEnd = 1e6
/
# the driver: an electron bunch
&Specie1
 Beam = 1
 Distribution = 1
 Density = 0.5
 RadiusX = 1.4142135
 Radius = 0.70710678
 x0 = 8.
 y0 = 0.25
 z0 = -0.5
 Begin = -100.
 End = 100.
 Polarity = -1.
 Px0 = 1000.
 P_perCell = 4
 MassAE = 0.000544
/
"""


def f32(x):
    return float(np.float32(x))


def test_nlist_reading_rules():
    text = ("#&Domain commented out\n"
            "junk line & Domain\n"
            " a = 1 2\n"
            "#b = 5\n"
            " c=2.5\n"
            " a = 3\n"
            " note = x/y ends here\n"
            " d = 4\n"
            "&Domain\n e = 6\n/\n")
    raw = read_nlist(text, "Domain")
    # the block starts at the first line holding '&Domain', ends at a '/'
    # a later "a" overrides; the line holding '/' closes the block unread
    assert raw == {"a": "3", "c": "2.5"}
    assert read_nlist(text, "Missing") is None
    with pytest.raises(DeckError):
        read_nlist("&Domain\n Hx = 1\n", "Domain")   # no closing '/'


def test_number_reading():
    from tertia_amd.deck import _as_double, _as_int
    assert _as_int("2.5", "k") == 2 and _as_int("-7x", "k") == -7
    assert _as_double("0.01", "k") == float(np.float32(0.01)) != 0.01   # %g into a float
    assert _as_double("1e-3,", "k") == float(np.float32(1e-3))
    with pytest.raises(DeckError):
        _as_int("x", "k")
    with pytest.raises(DeckError):
        _as_double("abc", "k")


def test_deck_to_params():
    p = load_deck(DECK)
    hx, hy = f32(0.05), f32(0.1875)
    assert p.hx == hx and p.nx_split == 1 and p.n_iter == 2
    # l_Xsize = long(Xlength/Hx)+1 layers, l_Ysize = long(Ylength/Hy)+1 cells
    assert p.nslices == int(f32(9.99) / hx) + 1 == 200
    assert p.ny == p.nz == int(f32(11.9) / hy) + 1 == 64
    assert p.ly == p.ny * hy and p.lz == p.nz * hy
    assert p.dens == 1.0 and p.ppc_y * p.ppc_z == 4 and p.jitter == 0.0
    assert p.shape == QSS_SHAPE_NGP_REF and p.fix_subcycle == 0
    (b,) = p.bunches
    assert b.nb == f32(0.5)
    assert b.sigma_xi == f32(1.4142135) / math.sqrt(2.0)
    assert b.sigma_r == f32(0.70710678) / math.sqrt(2.0)       # RadiusY = RadiusZ = Radius
    assert b.xi0 == p.nslices * hx - f32(8.0)                   # layer i at xi = (l_Mx - i) hx
    assert (b.y0, b.z0, b.polarity) == (f32(0.25), f32(-0.5), -1.0)
    assert abs(b.vx - 1000.0 / math.sqrt(1.0 + 1e6)) < 1e-15
    assert load_deck(DECK, shape=QSS_SHAPE_CIC).shape == QSS_SHAPE_CIC


@pytest.mark.parametrize("edit,key", [
    ("Distribution = 1\n Density = 0.5", "Distribution"),
    ("RadiusX = 1.4142135\n RadiusY = 0.3", "RadiusY"),
    ("Px0 = 1000.\n Delta = 0.1", "Delta"),
    ("Begin = -100.", "Begin"),
    ("Beam = 1", "Beam"),
])
def test_unsupported_settings_refused(edit, key):
    bad = {"Distribution": DECK.replace("Distribution = 1\n Density = 0.5",
                                        "Distribution = 2\n Density = 0.5"),
           "RadiusY": DECK.replace("RadiusX = 1.4142135", "RadiusX = 1.4142135\n RadiusY = 0.3"),
           "Delta": DECK.replace("Px0 = 1000.", "Px0 = 1000.\n Delta = 0.1"),
           "Begin": DECK.replace("Begin = -100.", "Begin = 5."),
           "Beam": DECK.replace("Beam = 1", "Beam = 0")}[key]
    with pytest.raises(DeckError) as e:
        load_deck(bad)
    assert key in str(e.value) or "ion" in str(e.value)


def test_plasma_must_be_uniform_and_2x2():
    with pytest.raises(DeckError):
        load_deck(DECK.replace("P_perCell = 4\n End", "P_perCell = 9\n End"))
    with pytest.raises(DeckError):
        load_deck(DECK.replace(" End = 1e6\n PlateauEnd = 1e6\n", ""))   # plateau ends at 0


def test_deck_to_driver():
    """Push parameters of the evolving driver: &Domain Ts / B*External /
    Wavelength, &Synchrotron SynMin, q2m = Polarity/(MassAE*1836) of the
    beam species (elements.cpp:141-185), all read as floats."""
    from tertia_amd.deck import load_driver
    d = load_driver(DECK)   # no Ts, no &Synchrotron: reference defaults
    assert d.ts == 1.0 and d.syn_min == 100.0 and d.wavelength == 1e-3
    assert list(d.b_ext) == [0.0, 0.0, 0.0]
    assert d.p0 == [(1000.0, 0.0, 0.0)]
    assert d.q2m == [f32(-1.0) / (f32(0.000544) * 1836)]
    deck = DECK.replace(" Nspecies = 2\n", " Nspecies = 2\n Ts = 2.5\n ByExternal = 0.1\n"
                        " Wavelength = 0.0008\n") + "&Synchrotron\n SynMin = 1e9\n/\n"
    d = load_driver(deck)
    assert d.ts == 2.5 and d.syn_min == f32(1e9) and d.wavelength == f32(0.0008)
    assert list(d.b_ext) == [0.0, f32(0.1), 0.0]
    with pytest.raises(DeckError):
        load_driver(DECK.replace(" Beam = 1\n", " Beam = 1\n Type = 0\n"))


def test_deck_external_b_reaches_the_plasma():
    """&Domain BxExternal..BzExternal (read as floats) are the plasma push's
    bXext..bZext (movepart_splitlayer.cpp:867-869; sf_DimFields = 1)."""
    assert list(load_deck(DECK).b_ext) == [0.0, 0.0, 0.0]
    deck = DECK.replace(" Nspecies = 2\n", " Nspecies = 2\n BxExternal = 0.3\n"
                        " BzExternal = -0.05\n")
    p = load_deck(deck)
    assert list(p.b_ext) == [f32(0.3), 0.0, f32(-0.05)]
    c = p.to_c()
    assert list(c.b_ext) == [f32(0.3), 0.0, f32(-0.05)]

