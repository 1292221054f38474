"""Parameters of the quasi-static xi-slice sweep (ctypes mirror of ``qss_params``).

The fields correspond to the reference's input-deck keys (``&Domain``:
``Hx, NxSplit, nIter, Hy, Hz, Ylength, Zlength``, domain.cpp:42-79;
``&Electrons``/``&Specie<i>``: ``Density, P_perCell, Beam, x0/y0/z0,
RadiusX/Y/Z``, elements.cpp:11-67,141-146).  The build uses exactly ``ny``
cells over ``ly`` (the reference makes ``long(L/h)+1``, domain.cpp:141-145).

Presets (the five workloads the sweep is measured and checked on; uniform
plasma n0=1, 2x2 particles per cell, immobile ions, frozen bi-Gaussian
drivers, CIC shape, 2 corrector iterations): ``linear_wake_64`` (64^2, 200
layers), ``blowout_256`` (256^2, 1024 layers), ``two_bunch_512`` (512^2,
2048 layers, driver + witness), ``parameter_scan`` (64 wakes on 128^2 ..
384^2) and ``profile_gradient_256`` (256^2, binned driver profile).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

QSS_KIND_C = 0
QSS_KIND_S = 1
QSS_SHAPE_NGP_REF = 0
QSS_SHAPE_CIC = 1
QSS_PREC_FP64 = 0
QSS_PREC_MIXED32 = 1
QSS_PREC_MIXED_BF16X3 = 2
QSS_PREC_MIXED_BF16X6 = 3
PRECISIONS = {"fp64": QSS_PREC_FP64, "mixed32": QSS_PREC_MIXED32,
              "bf16x3": QSS_PREC_MIXED_BF16X3, "bf16x6": QSS_PREC_MIXED_BF16X6}
QSS_MAX_BUNCH = 4

# field ids (enum qss_field)
FIELDS = ["Ex", "Ey", "Ez", "Bx", "By", "Bz", "Jx", "Jy", "Jz", "Rho",
          "RhoBeam", "JxBeam", "Psi"]
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}
QSS_NFIELD = len(FIELDS)

# spectral ids (order of the fft_of_* state, movefieldshydrolinlayersplit.cpp:267-313)
SPECTRAL = ["Rho", "Ex", "Ey", "Ez", "Bx", "By", "Bz", "Jx", "Jy", "Jz",
            "RhoP", "JxP", "JyP", "JzP", "ExP", "EyP", "EzP",
            "RhoBeam", "JxBeam", "RhoBeamP", "JxBeamP"]
SPECTRAL_ID = {n: i for i, n in enumerate(SPECTRAL)}

# transform kinds (z, y) per field, movefieldshydrolinlayersplit.cpp:389-427
FIELD_KINDS = {
    "Rho": (0, 0), "Ex": (0, 0), "Jx": (0, 0), "RhoBeam": (0, 0), "JxBeam": (0, 0),
    "Psi": (0, 0),
    "Ey": (0, 1), "Jy": (0, 1), "Bz": (0, 1),
    "Ez": (1, 0), "Jz": (1, 0), "By": (1, 0),
    "Bx": (1, 1),
}


class QssBunch(ctypes.Structure):
    _fields_ = [
        ("nb", ctypes.c_double),
        ("sigma_r", ctypes.c_double),
        ("sigma_xi", ctypes.c_double),
        ("xi0", ctypes.c_double),
        ("y0", ctypes.c_double),
        ("z0", ctypes.c_double),
        ("polarity", ctypes.c_double),
        ("vx", ctypes.c_double),
        ("noise", ctypes.c_double),
    ]


class QssParams(ctypes.Structure):
    _fields_ = [
        ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("ly", ctypes.c_double), ("lz", ctypes.c_double),
        ("hx", ctypes.c_double),
        ("nx_split", ctypes.c_int32),
        ("nslices", ctypes.c_int32),
        ("n_iter", ctypes.c_int32),
        ("dens", ctypes.c_double),
        ("ppc_y", ctypes.c_int32), ("ppc_z", ctypes.c_int32),
        ("jitter", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("shape", ctypes.c_int32),
        ("fix_subcycle", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("nbunch", ctypes.c_int32),
        ("bunch", QssBunch * QSS_MAX_BUNCH),
        ("nbins", ctypes.c_int32),
        ("bin_xi0", ctypes.c_double), ("bin_xi1", ctypes.c_double),
        ("bins", ctypes.POINTER(ctypes.c_double)),
        ("nwake", ctypes.c_int32),
        ("wake_bunch", ctypes.POINTER(QssBunch)),
        ("b_ext", ctypes.c_double * 3),
    ]


@dataclass
class Bunch:
    nb: float
    sigma_r: float
    sigma_xi: float
    xi0: float
    y0: float = 0.0
    z0: float = 0.0
    polarity: float = -1.0
    vx: float = 1.0
    noise: float = 0.0


@dataclass
class Params:
    ny: int = 64
    nz: int = 64
    ly: float = 12.0
    lz: float = 12.0
    hx: float = 0.05
    nx_split: int = 1
    nslices: int = 200
    n_iter: int = 2
    dens: float = 1.0
    ppc_y: int = 2
    ppc_z: int = 2
    jitter: float = 1e-3
    seed: int = 12345
    shape: int = QSS_SHAPE_CIC
    fix_subcycle: int = 0
    precision: int = QSS_PREC_FP64
    bunches: List[Bunch] = field(default_factory=list)
    bins: Optional[Sequence[float]] = None
    bin_xi0: float = 0.0
    bin_xi1: float = 0.0
    # batch of independent wakes on this grid: per-wake bunch lists (all the
    # same length as `bunches`); None = one wake
    wakes: Optional[List[List[Bunch]]] = None
    # uniform external B on the plasma (&Domain BxExternal..BzExternal)
    b_ext: Sequence[float] = (0.0, 0.0, 0.0)

    @property
    def n_particles(self) -> int:
        return self.ny * self.nz * self.ppc_y * self.ppc_z

    @property
    def nwake(self) -> int:
        return len(self.wakes) if self.wakes else 1

    def to_c(self) -> QssParams:
        if len(self.bunches) > QSS_MAX_BUNCH:
            raise ValueError("at most %d bunches" % QSS_MAX_BUNCH)
        c = QssParams()
        for name in ("ny", "nz", "ly", "lz", "hx", "nx_split", "nslices", "n_iter",
                     "dens", "ppc_y", "ppc_z", "jitter", "seed", "shape",
                     "fix_subcycle", "precision"):
            setattr(c, name, getattr(self, name))
        c.nbunch = len(self.bunches)
        for i, b in enumerate(self.bunches):
            for name in ("nb", "sigma_r", "sigma_xi", "xi0", "y0", "z0", "polarity",
                         "vx", "noise"):
                setattr(c.bunch[i], name, getattr(b, name))
        if self.bins is not None:
            arr = (ctypes.c_double * len(self.bins))(*[float(v) for v in self.bins])
            c.nbins = len(self.bins)
            c.bin_xi0 = self.bin_xi0
            c.bin_xi1 = self.bin_xi1
            c.bins = ctypes.cast(arr, ctypes.POINTER(ctypes.c_double))
            c._bins_keepalive = arr  # keep the buffer alive with the struct
        else:
            c.nbins = 0
            c.bins = None
        for a in range(3):
            c.b_ext[a] = float(self.b_ext[a])
        c.nwake = self.nwake
        if self.wakes:
            nbunch = len(self.bunches)
            if any(len(w) != nbunch for w in self.wakes):
                raise ValueError("every wake needs %d bunches" % nbunch)
            wb = (QssBunch * max(1, len(self.wakes) * nbunch))()
            for w, bl in enumerate(self.wakes):
                for i, b in enumerate(bl):
                    for name in ("nb", "sigma_r", "sigma_xi", "xi0", "y0", "z0", "polarity",
                                 "vx", "noise"):
                        setattr(wb[w * nbunch + i], name, getattr(b, name))
            c.wake_bunch = ctypes.cast(wb, ctypes.POINTER(QssBunch))
            c._wb_keepalive = wb
        else:
            c.wake_bunch = None
        return c


class QssDriver(ctypes.Structure):
    _fields_ = [
        ("ts", ctypes.c_double),
        ("b_ext", ctypes.c_double * 3),
        ("syn_min", ctypes.c_double),
        ("wavelength", ctypes.c_double),
        ("p0", (ctypes.c_double * 3) * QSS_MAX_BUNCH),
        ("q2m", ctypes.c_double * QSS_MAX_BUNCH),
    ]


@dataclass
class Driver:
    """Push parameters of an evolving driver (``qss_driver``): the time step
    ``Ts``, external B, the radiation-reaction threshold ``SynMin`` and
    ``Wavelength`` (&Domain / &Synchrotron keys, domain.cpp:50-74,
    synchrotron.cpp:13), and per bunch the seeded momentum and q2m
    (&Specie<i> ``Px0/Py0/Pz0``, q2m = Polarity/(MassAE*1836),
    elements.cpp:141-185)."""
    ts: float = 1.0
    p0: Sequence[Sequence[float]] = ((0.0, 0.0, 0.0),)
    q2m: Sequence[float] = (-1.0,)
    b_ext: Sequence[float] = (0.0, 0.0, 0.0)
    syn_min: float = 1e2
    wavelength: float = 1e-3

    def to_c(self) -> QssDriver:
        if len(self.p0) > QSS_MAX_BUNCH or len(self.q2m) > QSS_MAX_BUNCH:
            raise ValueError("at most %d bunches" % QSS_MAX_BUNCH)
        c = QssDriver()
        c.ts = self.ts
        c.syn_min = self.syn_min
        c.wavelength = self.wavelength
        for a in range(3):
            c.b_ext[a] = self.b_ext[a]
        for b, mom in enumerate(self.p0):
            for a in range(3):
                c.p0[b][a] = mom[a]
        for b, q in enumerate(self.q2m):
            c.q2m[b] = q
        return c


def linear_wake_64(**kw) -> Params:
    """12x12 k_p^-1 on 64x64, 200 slices dxi=0.05, driver sr=0.5 sxi=1 nb=0.5 at xi=2."""
    p = Params(ny=64, nz=64, ly=12.0, lz=12.0, hx=0.05, nslices=200,
               bunches=[Bunch(nb=0.5, sigma_r=0.5, sigma_xi=1.0, xi0=2.0)])
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def blowout_256(**kw) -> Params:
    """16x16 on 256x256, 1024 slices dxi=0.02, driver sr=0.3 sxi=0.8 nb=5, +-2% noise."""
    p = Params(ny=256, nz=256, ly=16.0, lz=16.0, hx=0.02, nslices=1024,
               bunches=[Bunch(nb=5.0, sigma_r=0.3, sigma_xi=0.8, xi0=2.0, noise=0.02)])
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def two_bunch_512(**kw) -> Params:
    """20x20 on 512x512, 2048 slices dxi=0.01; driver nb=4 sr=0.25 sxi=0.7 at
    xi=2.1 plus witness nb=8 sr=0.1 sxi=0.1 placed 4.5 behind it."""
    xd = 2.1
    p = Params(ny=512, nz=512, ly=20.0, lz=20.0, hx=0.01, nslices=2048,
               bunches=[Bunch(nb=4.0, sigma_r=0.25, sigma_xi=0.7, xi0=xd),
                        Bunch(nb=8.0, sigma_r=0.1, sigma_xi=0.1, xi0=xd + 4.5)])
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def parameter_scan(seed: int = 20261202, per_size: int = 16, nslices: int = 512,
            sizes=(128, 192, 256, 384)) -> List[Params]:
    """Parameter scan: per_size independent wakes on each of the transverse grids
    `sizes` (cell size 1/16 k_p^-1, so 8, 12, 16, 24 k_p^-1 boxes), 4 ppc,
    nslices layers of dxi = 0.02.  Each wake's driver has sigma_r ~ U[0.2, 0.8],
    sigma_xi ~ U[0.5, 1.5], nb ~ U[0.5, 8] drawn from `seed`, centred
    3 sigma_xi behind the front.  One batched Params per grid size."""
    import random
    rng = random.Random(seed)
    groups = []
    for n in sizes:
        wakes = []
        for _ in range(per_size):
            sr = rng.uniform(0.2, 0.8)
            sx = rng.uniform(0.5, 1.5)
            nb = rng.uniform(0.5, 8.0)
            wakes.append([Bunch(nb=nb, sigma_r=sr, sigma_xi=sx, xi0=3.0 * sx)])
        groups.append(Params(ny=n, nz=n, ly=n / 16.0, lz=n / 16.0, hx=0.02, nslices=nslices,
                             bunches=list(wakes[0]), wakes=wakes))
    return groups


def profile_gradient_256(**kw) -> Params:
    """256x256 over 16x16, 1024 slices dxi=0.01, FP64; driver longitudinal profile
    = 256 piecewise-constant bins over xi in [0,4] (initial Gaussian sxi=0.8,
    nb=3 centred at 2), sr=0.3."""
    nbins = 256
    x0, x1 = 0.0, 4.0
    w = (x1 - x0) / nbins
    bins = [3.0 * math.exp(-((x0 + (i + 0.5) * w - 2.0) ** 2) / (2 * 0.8 ** 2))
            for i in range(nbins)]
    p = Params(ny=256, nz=256, ly=16.0, lz=16.0, hx=0.01, nslices=1024,
               bunches=[Bunch(nb=3.0, sigma_r=0.3, sigma_xi=0.8, xi0=2.0)],
               bins=bins, bin_xi0=x0, bin_xi1=x1)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def small(ny=32, nz=32, nslices=20, **kw) -> Params:
    """A scaled-down linear-wake case that the oracle finishes in well under a second."""
    p = Params(ny=ny, nz=nz, ly=12.0 * ny / 64, lz=12.0 * nz / 64, hx=0.05,
               nslices=nslices,
               bunches=[Bunch(nb=0.5, sigma_r=0.5, sigma_xi=0.4, xi0=0.6)])
    for k, v in kw.items():
        setattr(p, k, v)
    return p
