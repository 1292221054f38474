"""ctypes wrapper of the FP64 CPU oracle (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py``.  It is the checker, never the
thing measured or shipped; the product path (``tertia_amd``) never imports it.

Every function of ``qss_oracle.cpp`` restates a reference routine; the
file:line it follows is given at each function there.  Pinning (each by a
check that does not restate the oracle's formula):

* transforms: scipy type-4 fixtures, brute force, the involution T.T = 2N.I
  (tests/test_oracle_transforms.py);
* Guess and Iterate, every output field: the continuous equations they
  discretise (continuity, Gauss, the cold-fluid current, Ampere, the psi
  Poisson equation), checked in real space with scipy-built derivatives on
  random planes, plus the predictor's plasma rotation
  (tests/test_oracle_fields.py, tests/test_oracle_spectral.py);
* push: Boris invariants, closed-form kicks, exact bilinear gather, deposit
  conservation (tests/test_oracle_push.py);
* frozen driver planes: charge, centroid and width of the analytic
  bi-Gaussian and of the binned profile (tests/test_oracle_beam.py);
* the whole sweep: the psi/Ex identity (tests/test_oracle_sweep.py) and
  second-order convergence to 3-D linear wake theory
  (tests/test_oracle_linear_wake.py);
* tests/test_oracle_mutants.py plants plausible mistakes, one at a time, in copies of the
  source and checks that these pins catch each of them.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from tertia_amd.params import FIELD_ID, SPECTRAL_ID, Driver, Params, QssDriver, QssParams

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []),
                       check=True)
    return _LIB_PATH


def _configure(L):
    """Declare the argument types of every exported oracle function on L."""
    L.oracle_create.restype = ctypes.c_void_p
    L.oracle_create.argtypes = [ctypes.POINTER(QssParams), ctypes.c_int,
                                ctypes.c_int, ctypes.c_uint32]
    L.oracle_destroy.argtypes = [ctypes.c_void_p]
    for name in ("oracle_reset", "oracle_n_particles"):
        getattr(L, name).argtypes = [ctypes.c_void_p]
    L.oracle_layer_step.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
    L.oracle_sweep_range.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
    L.oracle_sweep.argtypes = [ctypes.c_void_p, _dp]
    L.oracle_beam_layer.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.oracle_beam_plane.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
    L.oracle_guess.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.oracle_push_deposit.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.oracle_iterate.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
    L.oracle_copy_c_to_p.argtypes = [ctypes.c_void_p]
    L.oracle_get_layer_field.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
    L.oracle_set_layer_field.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
    L.oracle_get_spectral.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
    L.oracle_set_spectral.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
    L.oracle_get_history.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
    L.oracle_get_particles.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp,
                                       _i64p, _i32p, _i32p, _dp, _dp]
    L.oracle_set_particles.argtypes = [ctypes.c_void_p, ctypes.c_long, _i32p, _i32p,
                                       _dp, _dp, _dp, _dp, _dp, _i64p]
    L.oracle_set_history.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
    L.oracle_driver_seed.argtypes = [ctypes.c_void_p, ctypes.POINTER(QssDriver), _i64p]
    L.oracle_driver_push.argtypes = [ctypes.c_void_p]
    L.oracle_time_step.argtypes = [ctypes.c_void_p, _dp]
    L.oracle_driver_count.restype = ctypes.c_int64
    L.oracle_driver_count.argtypes = [ctypes.c_void_p]
    L.oracle_driver_get.argtypes = [ctypes.c_void_p, _i32p, _dp, _dp, _dp, _i32p]
    L.oracle_driver_set.argtypes = [ctypes.c_void_p, ctypes.c_int64, _i32p, _dp, _dp,
                                    _dp, _i32p]
    L.oracle_driver_density.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
    L.oracle_layer_record_doubles.restype = ctypes.c_long
    L.oracle_layer_record_doubles.argtypes = [ctypes.c_void_p]
    L.oracle_pack_layer_record.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
    L.oracle_unpack_layer_record.argtypes = [ctypes.c_void_p, _dp, ctypes.c_long]
    L.oracle_transform1d.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int]
    L.oracle_transform2d.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, _dp, _dp, ctypes.c_double,
                                     ctypes.c_int, ctypes.c_int]
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _configure(ctypes.CDLL(_LIB_PATH))
    return _lib


def load_library(path: str):
    """The oracle's functions from another build of its source: the tests that
    plant a mistake in a copy of qss_oracle.cpp load the copy with this and
    check that the pins fail on it."""
    return _configure(ctypes.CDLL(path))


def _p(a):
    return a.ctypes.data_as(_dp)


def transform1d(x, kind, use_fft=False):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    lib().oracle_transform1d(x.size, int(kind), _p(x), _p(out), int(use_fft))
    return out


def transform2d(x, kind_z, kind_y, scale=1.0, use_fft=False, nthreads=1):
    """x: (nz, ny) array; returns scale * T_kz . x . T_ky (FFTW r2r 2-D)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    nz, ny = x.shape
    out = np.empty_like(x)
    lib().oracle_transform2d(ny, nz, int(kind_z), int(kind_y), _p(x), _p(out),
                             float(scale), int(use_fft), int(nthreads))
    return out


class Oracle:
    """FP64 CPU restatement of Mesh::MoveAllSplitLayers and its pieces."""

    def __init__(self, params: Params, use_fft: bool = False, nthreads: int = 1,
                 history=("Ex", "Ey", "Ez", "Bx", "By", "Bz", "Jx", "Jy", "Jz",
                          "Rho", "RhoBeam", "JxBeam", "Psi"), library=None):
        self._L = library if library is not None else lib()
        self.params = params
        self._c = params.to_c()
        mask = 0
        for f in history or ():
            mask |= 1 << FIELD_ID[f]
        self.h = self._L.oracle_create(ctypes.byref(self._c), int(use_fft), int(nthreads),
                                     mask)
        if not self.h:
            raise ValueError("oracle_create rejected the parameters")
        self.ny, self.nz = params.ny, params.nz

    def close(self):
        if self.h:
            self._L.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n_particles(self):
        return self._L.oracle_n_particles(self.h)

    def _plane(self):
        return np.zeros((self.nz, self.ny), dtype=np.float64)

    # -- sweep ---------------------------------------------------------
    def reset(self):
        self._L.oracle_reset(self.h)

    def sweep(self):
        n = self.params.nslices * self.params.nx_split
        res = np.zeros(n)
        st = self._L.oracle_sweep(self.h, _p(res))
        return st, res

    def sweep_range(self, first, count):
        res = np.zeros(count * self.params.nx_split)
        st = self._L.oracle_sweep_range(self.h, first, count, _p(res))
        return st, res

    def layer_step(self, slice_, isplit=0):
        r = np.zeros(1)
        st = self._L.oracle_layer_step(self.h, slice_, isplit, _p(r))
        return st, float(r[0])

    # -- pieces ----------------------------------------------------------
    def beam_layer(self, slice_, isplit=0):
        return self._L.oracle_beam_layer(self.h, slice_, isplit)

    def beam_plane(self, layer):
        r, j = self._plane(), self._plane()
        self._L.oracle_beam_plane(self.h, layer, _p(r), _p(j))
        return r, j

    def guess(self, first_layer):
        return self._L.oracle_guess(self.h, int(first_layer))

    def push_deposit(self, full_step):
        return self._L.oracle_push_deposit(self.h, int(full_step))

    def iterate(self, final_pass):
        r = np.zeros(1)
        self._L.oracle_iterate(self.h, int(final_pass), _p(r))
        return float(r[0])

    def copy_c_to_p(self):
        return self._L.oracle_copy_c_to_p(self.h)

    # -- data ------------------------------------------------------------
    def get_field(self, name, which=0):
        out = self._plane()
        self._L.oracle_get_layer_field(self.h, which, FIELD_ID[name], _p(out))
        return out

    def set_field(self, name, arr, which=0):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.nz, self.ny)
        self._L.oracle_set_layer_field(self.h, which, FIELD_ID[name], _p(a))

    def get_spectral(self, name):
        out = self._plane()
        self._L.oracle_get_spectral(self.h, SPECTRAL_ID[name], _p(out))
        return out

    def set_spectral(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.nz, self.ny)
        self._L.oracle_set_spectral(self.h, SPECTRAL_ID[name], _p(a))

    def history(self, name, slice_):
        out = self._plane()
        st = self._L.oracle_get_history(self.h, FIELD_ID[name], slice_, _p(out))
        if st != 0:
            raise ValueError("no history for %s" % name)
        return out

    def particles(self):
        n = self.n_particles
        d = {k: np.zeros(n) for k in ("y", "z", "px", "py", "pz", "fy", "fz")}
        d["id"] = np.zeros(n, dtype=np.int64)
        d["cj"] = np.zeros(n, dtype=np.int32)
        d["ck"] = np.zeros(n, dtype=np.int32)
        self._L.oracle_get_particles(self.h, _p(d["y"]), _p(d["z"]), _p(d["px"]),
                                   _p(d["py"]), _p(d["pz"]),
                                   d["id"].ctypes.data_as(_i64p),
                                   d["cj"].ctypes.data_as(_i32p),
                                   d["ck"].ctypes.data_as(_i32p),
                                   _p(d["fy"]), _p(d["fz"]))
        return d

    def set_particles(self, cj, ck, fy, fz, px, py, pz, pid):
        arrs = [np.ascontiguousarray(a, dtype=np.int32) for a in (cj, ck)]
        f = [np.ascontiguousarray(a, dtype=np.float64) for a in (fy, fz, px, py, pz)]
        pid = np.ascontiguousarray(pid, dtype=np.int64)
        self._L.oracle_set_particles(self.h, len(pid), arrs[0].ctypes.data_as(_i32p),
                                   arrs[1].ctypes.data_as(_i32p), _p(f[0]), _p(f[1]),
                                   _p(f[2]), _p(f[3]), _p(f[4]),
                                   pid.ctypes.data_as(_i64p))

    def set_history(self, name, slice_, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.nz, self.ny)
        if self._L.oracle_set_history(self.h, FIELD_ID[name], slice_, _p(a)) != 0:
            raise ValueError("no history for %s" % name)

    # -- the reference's layer record (PackLayer / UnpackLayer) ----------
    def pack_layer_record(self, iLayer):
        buf = np.empty(self._L.oracle_layer_record_doubles(self.h), dtype=np.float64)
        self._L.oracle_pack_layer_record(self.h, int(iLayer), _p(buf))
        return buf

    def unpack_layer_record(self, rec):
        r = np.ascontiguousarray(rec, dtype=np.float64)
        if self._L.oracle_unpack_layer_record(self.h, _p(r), r.size) != 0:
            raise ValueError("not a layer record of this grid")

    # -- evolving driver -----------------------------------------------
    def driver_seed(self, driver: Driver) -> int:
        self._drv = driver.to_c()
        n = ctypes.c_int64(0)
        if self._L.oracle_driver_seed(self.h, ctypes.byref(self._drv), ctypes.byref(n)) != 0:
            raise ValueError("the driver needs the history of Ex..Bz")
        return n.value

    def driver_push(self):
        st = self._L.oracle_driver_push(self.h)
        if st != 0:
            raise ValueError("oracle_driver_push: status %d" % st)

    def time_step(self):
        res = np.zeros(self.params.nslices * self.params.nx_split)
        st = self._L.oracle_time_step(self.h, _p(res))
        if st != 0:
            raise ValueError("oracle_time_step: status %d" % st)
        return res

    def driver_particles(self):
        n = self._L.oracle_driver_count(self.h)
        d = {"cell": np.zeros((n, 3), dtype=np.int32), "frac": np.zeros((n, 3)),
             "mom": np.zeros((n, 3)), "weight": np.zeros(n),
             "bunch": np.zeros(n, dtype=np.int32)}
        self._L.oracle_driver_get(self.h, d["cell"].ctypes.data_as(_i32p), _p(d["frac"]),
                                _p(d["mom"]), _p(d["weight"]),
                                d["bunch"].ctypes.data_as(_i32p))
        return d

    def driver_set(self, cell, frac, mom, weight, bunch):
        cell = np.ascontiguousarray(cell, dtype=np.int32).reshape(-1, 3)
        frac = np.ascontiguousarray(frac, dtype=np.float64).reshape(-1, 3)
        mom = np.ascontiguousarray(mom, dtype=np.float64).reshape(-1, 3)
        weight = np.ascontiguousarray(weight, dtype=np.float64)
        bunch = np.ascontiguousarray(bunch, dtype=np.int32)
        if self._L.oracle_driver_set(self.h, len(weight), cell.ctypes.data_as(_i32p),
                                   _p(frac), _p(mom), _p(weight),
                                   bunch.ctypes.data_as(_i32p)) != 0:
            raise ValueError("seed the driver first")

    def driver_density(self, layer):
        rho, jx = self._plane(), self._plane()
        if self._L.oracle_driver_density(self.h, int(layer), _p(rho), _p(jx)) != 0:
            raise ValueError("no driver deposit for layer %d" % layer)
        return rho, jx
