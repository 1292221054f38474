"""Reference-style input decks (NList namelists) -> sweep parameters.

A deck is the text file the reference reads with ``NList::read``
(namelist.cpp:93-146): a block starts at the first line containing ``&Name``
and ends at the next line containing ``/``; lines whose first character is
``#`` are skipped; inside a block each line is split at blanks, tabs, ``=``
and newlines into ``key value``; unknown keys are ignored; a key given twice
keeps the later value; a later block of the same name is never read (every
list rewinds and takes the first match).
Integers are read with ``%d`` (``"2.5"`` gives 2); doubles are read with
``%g`` into a *float* and widened (namelist.cpp:128-133), so ``Hx = 0.01``
becomes 0.009999999776482582 — reproduced here, since it sets the layer
thickness the reference actually runs with.

The lists used by the xi sweep:

* ``&Domain`` (domain.cpp:42-79): ``Hx, Hy, Hz, Xlength, Ylength, Zlength,
  NxSplit`` (default 1), ``nIter`` (default 1), ``Nspecies`` (default 1),
  ``BxExternal, ByExternal, BzExternal`` (default 0): the uniform external B
  that the plasma push adds to the gathered field (bXext..bZext,
  movepart_splitlayer.cpp:867-869, 1271-1273; Cell::sf_DimFields = 1,
  domain.cpp:204).
  Grid sizing as the reference (domain.cpp:141-145, one rank):
  ``l_Xsize = long(Xlength/Hx) + 1`` layers, ``l_Ysize = long(Ylength/Hy) + 1``
  cells over ``l_Ysize * Hy`` (same for z).
* ``&Electrons`` (elements.cpp:11-67,88-113): ``Density`` (n0), ``P_perCell``,
  ``Distribution``; weight ``1/P_perCell``, charge -1.  The plasma density of
  the field solver is ``sum |Density * q2m|`` over the non-beam species
  (movefieldshydrolinlayersplit.cpp:244-252).
* ``&Specie<i>``, i = 1..Nspecies-1 (domain.cpp:165-174, elements.cpp:141-185)
  with ``Beam = 1`` and ``Distribution = 1`` (the Gaussian ellipsoid of
  plasma.cpp:350-377): ``Density, Radius, RadiusX/Y/Z, x0, y0, z0, Begin,
  End, Delta, Polarity, Px0/Py0/Pz0, P_perCell``.  The beam density is
  evaluated at ``xco = X(i) = i*Hx`` (mesh.cpp:108-115, mesh.h:246-249) and the
  sweep's frozen driver puts layer i at xi = (l_Mx - i)*hx, so
  ``xi0 = l_Mx*Hx - x0``; ``R = sqrt(2) sigma`` gives ``sigma = Radius/sqrt(2)``;
  ``y0, z0`` are already relative to the box centre (mesh.cpp:113-114); the
  frozen velocity is ``Px0 / sqrt(1 + |P0|^2)``.

What the sweep cannot represent is refused with a ``DeckError`` naming the
key, never silently approximated: plasma profiles other than a trapezoid that
is flat over the whole box,
mobile ion species, beam profiles other than the Gaussian ellipsoid, unequal
``RadiusY``/``RadiusZ``, a density modulation (``Delta``), a ``Begin``/``End``
cut inside the bunch, and particles-per-cell counts other than the 2x2
plasma lattice (mesh.cpp:455-459) and the 4-per-cell beam lattice
(mesh.cpp:151-155).
"""
from __future__ import annotations

import math
import os
from typing import Dict, List, Optional, Union

import numpy as np

from .params import QSS_SHAPE_NGP_REF, Bunch, Driver, Params


class DeckError(ValueError):
    """A deck the xi sweep cannot run as the reference would."""


def _tokens(line: str) -> List[str]:
    for ch in "=\t\n":
        line = line.replace(ch, " ")
    return line.split()


def read_nlist(text: str, name: str) -> Optional[Dict[str, str]]:
    """Raw ``key -> value`` strings of the first ``&name`` block (NList::read,
    namelist.cpp:93-146); None when the deck has no such block."""
    lines = text.splitlines(keepends=True)
    i = 0
    found = False
    while i < len(lines):
        ln = lines[i]
        i += 1
        if ln.startswith("#"):
            continue
        at = ln.find("&")
        if at >= 0:
            toks = ln[at + 1:].split()
            if toks and toks[0] == name:
                found = True
                break
    if not found:
        return None
    out: Dict[str, str] = {}
    while i < len(lines):
        ln = lines[i]
        i += 1
        if ln.startswith("#"):
            continue
        if "/" in ln:
            return out
        toks = [t for t in _tokens(ln) if t]
        if len(toks) >= 2:
            out[toks[0]] = toks[1]   # a repeated key: the later line is read last
    raise DeckError("&%s: premature end of file (no closing '/')" % name)


def _as_int(v: str, key: str) -> int:
    # sscanf("%d"): optional sign and leading digits
    s = v.strip()
    n = 0
    while n < len(s) and (s[n].isdigit() or (n == 0 and s[n] in "+-")):
        n += 1
    if n == 0 or s[:n] in "+-":
        raise DeckError("%s: cannot read an integer from %r" % (key, v))
    return int(s[:n])


def _as_double(v: str, key: str) -> float:
    # sscanf("%g") into a float, then widened to double
    s = v.strip()
    best = None
    for n in range(len(s), 0, -1):
        try:
            best = float(s[:n])
            break
        except ValueError:
            continue
    if best is None:
        raise DeckError("%s: cannot read a number from %r" % (key, v))
    return float(np.float32(best))


class _List:
    def __init__(self, raw: Dict[str, str], name: str):
        self.raw, self.name = raw, name

    def i(self, key: str, default: Optional[int] = None) -> int:
        if key in self.raw:
            return _as_int(self.raw[key], "&%s %s" % (self.name, key))
        if default is None:
            raise DeckError("&%s: %s is required" % (self.name, key))
        return default

    def d(self, key: str, default: Optional[float] = None) -> float:
        if key in self.raw:
            return _as_double(self.raw[key], "&%s %s" % (self.name, key))
        if default is None:
            raise DeckError("&%s: %s is required" % (self.name, key))
        return default


def load_deck(deck: Union[str, os.PathLike], shape: int = QSS_SHAPE_NGP_REF,
              n_iter: Optional[int] = None) -> Params:
    """Parameters of the xi sweep for a reference deck (a path or the deck
    text).  ``shape`` defaults to the reference's linked split path (NGP
    gather/deposit, QSS_SHAPE_NGP_REF); pass QSS_SHAPE_CIC for the CIC shape.
    ``n_iter`` overrides the deck's ``nIter``."""
    text = str(deck)
    if "\n" not in text and os.path.exists(text):
        with open(text) as f:
            text = f.read()
    raw = read_nlist(text, "Domain")
    if raw is None:
        raise DeckError("no &Domain list")
    dom = _List(raw, "Domain")
    hx, hy, hz = dom.d("Hx"), dom.d("Hy"), dom.d("Hz")
    xl, yl, zl = dom.d("Xlength"), dom.d("Ylength"), dom.d("Zlength")
    if hx <= 0 or xl <= hx or hy <= 0 or yl <= hy or hz <= 0 or zl <= hz:
        raise DeckError("Wrong step or domain size (domain.cpp:117-122)")
    nx_split = max(1, dom.i("NxSplit", 1))
    niter = max(0, dom.i("nIter", 1)) if n_iter is None else int(n_iter)
    nsp = dom.i("Nspecies", 1)
    nslices = int(xl / hx) + 1
    ny, nz = int(yl / hy) + 1, int(zl / hz) + 1

    raw = read_nlist(text, "Electrons")
    if raw is None:
        raise DeckError("no &Electrons list")
    el = _List(raw, "Electrons")
    if el.i("Distribution", 0) != 0:
        raise DeckError("&Electrons Distribution: only the uniform plasma (0) is supported")
    ppc = el.i("P_perCell", 0)
    if ppc != 4:
        raise DeckError("&Electrons P_perCell = %d: the sweep seeds the 2x2 lattice "
                        "(P_perCell = 4)" % ppc)
    n0 = el.d("Density")
    # the trapezoid of Distribution 0 (plasma.cpp:381-393) must be flat over
    # every layer, xco = 0 .. l_Mx*Hx (unset keys read as 0, namelist.h:21-24)
    box = nslices * hx
    if (el.d("Begin", 0.0) > 0.0 or el.d("PlateauBegin", 0.0) > 0.0 or
            el.d("PlateauEnd", 0.0) < box or el.d("End", 0.0) < box):
        raise DeckError("&Electrons Begin/PlateauBegin/PlateauEnd/End: the plasma must be "
                        "uniform over the whole box [0, %g]" % box)
    if n0 <= 0:
        raise DeckError("&Electrons Density must be positive")
    dens = abs(n0 * -1.0)   # sum |Density * q2m| over non-beam species, q2m = -1

    bunches: List[Bunch] = []
    for isp in range(1, nsp):
        nm = "Specie%d" % isp
        raw = read_nlist(text, nm)
        if raw is None:
            raise DeckError("Nspecies = %d but no &%s list" % (nsp, nm))
        sp = _List(raw, nm)
        if sp.i("Beam", 0) == 0:
            raise DeckError("&%s: mobile ion species are not part of the xi sweep "
                            "(the ions are an immobile background)" % nm)
        if sp.i("Distribution", 0) != 1:
            raise DeckError("&%s Distribution: only the Gaussian ellipsoid (1) is supported"
                            % nm)
        if sp.i("P_perCell", 0) != 4:
            raise DeckError("&%s P_perCell: the frozen driver seeds 4 per cell" % nm)
        if sp.d("Delta", 0.0) != 0.0:
            raise DeckError("&%s Delta: a longitudinal modulation is not supported" % nm)
        r = sp.d("Radius", 0.0)
        if r <= 0:
            r = 1.0                              # plasma.cpp:355-357
        rx = sp.d("RadiusX", 0.0) or r
        ry = sp.d("RadiusY", 0.0) or r
        rz = sp.d("RadiusZ", 0.0) or r
        if ry != rz:
            raise DeckError("&%s RadiusY != RadiusZ: the sweep's bunches are round" % nm)
        x0 = sp.d("x0", 0.0)
        begin, end = sp.d("Begin", 0.0), sp.d("End", 0.0)
        # the exp(-arg) profile is cut at arg > 16 (4 radii); a Begin/End
        # window inside that range would cut the bunch
        if begin > x0 - 4 * rx or end < x0 + 4 * rx:
            raise DeckError("&%s Begin/End cut the bunch (need Begin <= x0 - 4 RadiusX and "
                            "End >= x0 + 4 RadiusX)" % nm)
        px, py, pz = sp.d("Px0", 0.0), sp.d("Py0", 0.0), sp.d("Pz0", 0.0)
        vx = px / math.sqrt(1.0 + px * px + py * py + pz * pz)
        bunches.append(Bunch(nb=sp.d("Density"), sigma_r=ry / math.sqrt(2.0),
                             sigma_xi=rx / math.sqrt(2.0), xi0=nslices * hx - x0,
                             y0=sp.d("y0", 0.0), z0=sp.d("z0", 0.0),
                             polarity=sp.d("Polarity", 1.0), vx=vx))
    if len(bunches) > 4:
        raise DeckError("at most 4 beam species")
    b_ext = (dom.d("BxExternal", 0.0), dom.d("ByExternal", 0.0), dom.d("BzExternal", 0.0))
    return Params(ny=ny, nz=nz, ly=ny * hy, lz=nz * hz, hx=hx, nx_split=nx_split,
                  nslices=nslices, n_iter=niter, dens=dens, ppc_y=2, ppc_z=2, jitter=0.0,
                  shape=shape, fix_subcycle=0, bunches=bunches, b_ext=b_ext)


def load_driver(deck: Union[str, os.PathLike]) -> Driver:
    """Push parameters of the evolving driver (Mesh::MoveBeamParticles,
    movebeam.cpp:14-660) for a reference deck:

    * ``&Domain`` ``Ts`` (default 1), ``BxExternal/ByExternal/BzExternal``
      (default 0; Cell::sf_DimFields = 1, domain.cpp:204) and ``Wavelength``
      (default 0.001), domain.cpp:50-74;
    * ``&Synchrotron`` ``SynMin`` (default 100, synchrotron.cpp:13): the
      radiation-reaction threshold on gamma;
    * per beam ``&Specie<i>`` (in the order of :func:`load_deck`'s bunches):
      ``Px0, Py0, Pz0`` (elements.cpp:59) and q2m = Polarity/(MassAE*1836)
      (elements.cpp:141-185, the ion species a beam is seeded as by default,
      ``Type`` = 1, mesh.cpp:246-255).  Other species types are refused: the
      reference seeds them as Electron objects of species 0."""
    text = str(deck)
    if "\n" not in text and os.path.exists(text):
        with open(text) as f:
            text = f.read()
    raw = read_nlist(text, "Domain")
    if raw is None:
        raise DeckError("no &Domain list")
    dom = _List(raw, "Domain")
    raw = read_nlist(text, "Synchrotron")
    syn = _List(raw, "Synchrotron").d("SynMin", 1e2) if raw is not None else 1e2
    p0, q2m = [], []
    for isp in range(1, dom.i("Nspecies", 1)):
        nm = "Specie%d" % isp
        raw = read_nlist(text, nm)
        if raw is None:
            raise DeckError("no &%s list" % nm)
        sp = _List(raw, nm)
        if sp.i("Beam", 0) == 0:
            continue
        if sp.i("Type", 1) != 1:
            raise DeckError("&%s Type: only beam species seeded as ions (Type = 1) have "
                            "their own q2m" % nm)
        p0.append((sp.d("Px0", 0.0), sp.d("Py0", 0.0), sp.d("Pz0", 0.0)))
        q2m.append(sp.d("Polarity", 1.0) / (sp.d("MassAE", 1.0) * 1836))
    return Driver(ts=dom.d("Ts", 1.0), p0=p0 or [(0.0, 0.0, 0.0)], q2m=q2m or [0.0],
                  b_ext=(dom.d("BxExternal", 0.0), dom.d("ByExternal", 0.0),
                         dom.d("BzExternal", 0.0)),
                  syn_min=syn, wavelength=dom.d("Wavelength", 1e-3))
