"""Host-side mirror of the reference ``Mesh`` layer-step interface.

The reference drives the quasi-static sweep through member functions of its
single ``Mesh`` (mesh.h:51-84).  This class keeps those names and argument
meanings so that a caller of the reference finds the same call sequence, but
every call runs on the GPU through ``libqss.so``:

=============================================  =====================================
reference (file:line)                          here
=============================================  =====================================
Mesh::MoveAllSplitLayers   movepart_splitlayer.cpp:25    ``MoveAllSplitLayers``
Mesh::MoveSplitLayer       movepart_splitlayer.cpp:561   ``MoveSplitLayer``
beam C/P interpolation     movepart_splitlayer.cpp:330   ``InterpolateBeamLayers``
GuessFieldsHydroLinLayerSplit  movefieldshydrolinlayersplit.cpp:222  ``GuessFieldsHydroLinLayerSplit``
IterateFieldsHydroLinLayerSplit :1173 (returns Eerr)     ``IterateFieldsHydroLinLayerSplit``
MoveParticlesLayerSplit    movepart_splitlayer.cpp:781   ``MoveParticlesLayerSplit``
ClearCurrentsSplit/ClearRhoSplit mesh.cpp:692-708        fused into the deposit
ExchangeCurrentsSplit/ExchangeFieldsSplit :1659-1711    fused (periodic wrap)
=============================================  =====================================

Layer indices follow the reference: ``iLayer`` runs from ``l_Mx-1`` (first
layer behind the front) down to 0; ``slice = l_Mx-1-iLayer``.  Errors raise
:class:`QssError` where the reference calls ``exit()``.
"""
from __future__ import annotations

import ctypes
import os
from typing import Iterable, Optional

import numpy as np

from . import _lib
from ._lib import QssError, check
from .params import FIELD_ID, SPECTRAL_ID, Driver, Params

_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)


def _p(a):
    return a.ctypes.data_as(_dp)


class Mesh:
    """One quasi-static wake on one GPU (a ``qss_ctx``)."""

    def __init__(self, params: Params, device: int = 0,
                 history: Optional[Iterable[str]] = ("Ex", "Psi")):
        self.params = params
        self._c = params.to_c()
        mask = 0
        for f in history or ():
            mask |= 1 << FIELD_ID[f]
        self.L = _lib.load()
        h = ctypes.c_void_p()
        check(self.L.qss_create(ctypes.byref(self._c), device, mask, ctypes.byref(h)),
              "qss_create")
        self.h = h
        self.ny, self.nz = params.ny, params.nz
        self.l_Mx = params.nslices
        # bumped by every checkpointed sweep and every change of the profile
        # bins: an adjoint taken for an older forward would be that of the
        # newer one (the reverse sweep recomputes from the latest checkpoints)
        self.checkpoint_generation = 0

    # -- lifetime --------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.L.qss_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def n_particles(self) -> int:
        n = ctypes.c_int64()
        check(self.L.qss_info(self.h, ctypes.byref(n), None), "qss_info")
        return n.value

    @property
    def stream(self) -> int:
        return self.L.qss_stream(self.h)

    def _slice(self, iLayer: int) -> int:
        s = self.l_Mx - 1 - int(iLayer)
        if not 0 <= s < self.l_Mx:
            raise QssError(-1, "layer index", "iLayer %d outside [0, %d)" % (iLayer, self.l_Mx))
        return s

    # -- the sweep -------------------------------------------------------
    def MoveAllSplitLayers(self) -> np.ndarray:
        """Seed the front layer, sweep every layer; returns the final-pass
        residual of every layer step (movepart_splitlayer.cpp:25-558)."""
        nw = self.params.nwake
        res = np.zeros(self.l_Mx * self.params.nx_split * nw)
        check(self.L.qss_sweep(self.h, _p(res)), "qss_sweep")
        return res if nw == 1 else res.reshape(nw, -1)

    def sweep_range(self, first_slice: int, count: int) -> np.ndarray:
        """Sweep layers [first, first+count); residuals (per wake for a batch)."""
        nw = self.params.nwake
        res = np.zeros(count * self.params.nx_split * nw)
        check(self.L.qss_sweep_range(self.h, first_slice, count, _p(res)), "qss_sweep_range")
        return res if nw == 1 else res.reshape(nw, -1)

    def enable_graph(self, on: bool = True) -> None:
        """Replay layer-step pairs from a captured hipGraph in sweeps."""
        check(self.L.qss_enable_graph(self.h, int(on)), "qss_enable_graph")

    def select_wake(self, w: int) -> None:
        """Batched context: the wake that field/history/particle access reads."""
        check(self.L.qss_select_wake(self.h, int(w)), "qss_select_wake")

    def reset(self):
        check(self.L.qss_reset(self.h), "qss_reset")

    def MoveSplitLayer(self, iLayer: int, iSplit: int = 0) -> float:
        """One layer step (movepart_splitlayer.cpp:561-778), beam interpolation
        and the closing C->P copy included; returns the final Eerr."""
        r = ctypes.c_double()
        check(self.L.qss_layer_step(self.h, self._slice(iLayer), iSplit, ctypes.byref(r)),
              "qss_layer_step")
        return r.value

    # -- pieces of a layer step -------------------------------------------
    def InterpolateBeamLayers(self, iLayer: int, iSplit: int = 0):
        check(self.L.qss_beam_layer(self.h, self._slice(iLayer), iSplit), "qss_beam_layer")

    def GuessFieldsHydroLinLayerSplit(self, iLayer: int, iSplit: int = 0):
        check(self.L.qss_guess(self.h, int(iLayer == self.l_Mx - 1)), "qss_guess")

    def MoveParticlesLayerSplit(self, iLayer: int, iSplit: int, iFullStep: int,
                                part: float = 1.0):
        if part != 1.0:
            raise QssError(-1, "MoveParticlesLayerSplit", "only part=1 is used by the sweep")
        check(self.L.qss_push_deposit(self.h, int(iFullStep)), "qss_push_deposit")

    def IterateFieldsHydroLinLayerSplit(self, iLayer: int, iSplit: int, N_iter: int) -> float:
        r = ctypes.c_double()
        check(self.L.qss_iterate(self.h, int(N_iter < 0), ctypes.byref(r)), "qss_iterate")
        return r.value

    def ClearCurrentsSplit(self):
        """Fused into MoveParticlesLayerSplit (mesh.cpp:692-700)."""

    def ClearRhoSplit(self):
        """Fused into MoveParticlesLayerSplit (mesh.cpp:702-708)."""

    def ExchangeCurrentsSplit(self, iLayer: int = 0):
        """Periodic ghost fold, fused into the deposit (movepart_splitlayer.cpp:1685)."""

    def ExchangeFieldsSplit(self, iLayer: int = 0):
        """Periodic ghost copy, fused into the gather (movepart_splitlayer.cpp:1659)."""

    def CopyCtoP(self):
        check(self.L.qss_copy_c_to_p(self.h), "qss_copy_c_to_p")

    # -- data --------------------------------------------------------------
    def _plane(self):
        return np.zeros((self.nz, self.ny), dtype=np.float64)

    def get_field(self, name: str, layer: str = "C") -> np.ndarray:
        out = self._plane()
        check(self.L.qss_get_layer_field(self.h, 0 if layer == "C" else 1, FIELD_ID[name],
                                         _p(out)), "qss_get_layer_field")
        return out

    def set_field(self, name: str, arr, layer: str = "C"):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.nz, self.ny)
        check(self.L.qss_set_layer_field(self.h, 0 if layer == "C" else 1, FIELD_ID[name],
                                         _p(a)), "qss_set_layer_field")

    def get_spectral(self, name: str) -> np.ndarray:
        out = self._plane()
        check(self.L.qss_get_spectral(self.h, SPECTRAL_ID[name], _p(out)), "qss_get_spectral")
        return out

    def set_spectral(self, name: str, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.nz, self.ny)
        check(self.L.qss_set_spectral(self.h, SPECTRAL_ID[name], _p(a)), "qss_set_spectral")

    def history(self, name: str, slice_: int) -> np.ndarray:
        out = self._plane()
        check(self.L.qss_get_history(self.h, FIELD_ID[name], slice_, _p(out)),
              "qss_get_history")
        return out

    def set_history(self, name: str, slice_: int, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.nz, self.ny)
        check(self.L.qss_set_history(self.h, FIELD_ID[name], slice_, _p(a)),
              "qss_set_history")

    # -- layer hand-off between xi chunks (para.cpp:94-412) -------------------
    def pack_layer(self) -> np.ndarray:
        """The state a sweep continues from (after the last swept layer) as
        a flat byte buffer: the wire format of a chunk boundary."""
        n = ctypes.c_size_t()
        check(self.L.qss_layer_state_bytes(self.h, ctypes.byref(n)), "qss_layer_state_bytes")
        buf = np.empty(n.value, dtype=np.uint8)
        check(self.L.qss_pack_layer(self.h, buf.ctypes.data_as(ctypes.c_void_p)),
              "qss_pack_layer")
        return buf

    def unpack_layer(self, buf: np.ndarray):
        buf = np.ascontiguousarray(buf, dtype=np.uint8)
        check(self.L.qss_unpack_layer(self.h, buf.ctypes.data_as(ctypes.c_void_p), buf.size),
              "qss_unpack_layer")

    def PackLayer(self, iLayer: int) -> np.ndarray:
        """The reference's layer record (PackLayer, para.cpp:94-160) of the
        last swept layer, written as reference layer iLayer: 12 planes of
        Ny*Nz doubles, then 12 attribute arrays of Np particles."""
        n = ctypes.c_size_t()
        check(self.L.qss_layer_record_doubles(self.h, ctypes.byref(n)),
              "qss_layer_record_doubles")
        buf = np.empty(n.value, dtype=np.float64)
        check(self.L.qss_pack_layer_record(self.h, int(iLayer), _p(buf)), "qss_pack_layer_record")
        return buf

    def UnpackLayer(self, record: np.ndarray):
        """Continue from a reference layer record (UnpackLayer, para.cpp:167-280):
        Np follows from its length as in ReceiveLayer (para.cpp:385)."""
        r = np.ascontiguousarray(record, dtype=np.float64)
        check(self.L.qss_unpack_layer_record(self.h, _p(r), r.size), "qss_unpack_layer_record")

    # -- HDF5 output and restart: Mesh::Save / Mesh::Load --------------------
    def Save(self, isave: int, directory: str = ".", phase: float = 0.0, shift: int = 0) -> None:
        """Mesh::Save (save.cpp:178-835): writes vs<isave>_3d_fields.h5 (the
        swept wake as [Mz][My][Mx] fields ex..jz, dens, Dens<sort>, with the
        axes, meshinfo and phase) and vs<isave>_3d_particles.h5 (the evolving
        driver's particles) into `directory`.  Needs the history of ex..jz
        (and of rhobeam with the frozen driver)."""
        check(self.L.qss_save(self.h, directory.encode(), int(isave), float(phase), int(shift)),
              "qss_save")

    def Load(self, isave: int, directory: str = ".") -> float:
        """Mesh::Load (load.cpp:285-1080): the history of ex..jz that this
        mesh keeps, and a seeded evolving driver's particles, from the files
        Save wrote; returns the phase."""
        ph = ctypes.c_double()
        check(self.L.qss_load(self.h, directory.encode(), int(isave), ctypes.byref(ph)),
              "qss_load")
        if getattr(self, "_driver_n", None) is not None:   # particles were replaced
            from . import h5
            path = os.path.join(directory, "vs%03d_3d_particles.h5" % int(isave))
            have = set(h5.names(path))
            self._driver_n = sum(h5.info(path, "Ions_of_sort_%02d_at_PE_0" % s)[1][0]
                                 for s in range(1, len(self.params.bunches) + 1)
                                 if "Ions_of_sort_%02d_at_PE_0" % s in have)
        return ph.value

    # -- evolving driver: Domain::Step (step.cpp:63,105) ---------------------
    def SeedBeamParticles(self, driver: Driver) -> int:
        """Mesh::SeedBeamParticles (mesh.cpp:43-64,108-276) of every bunch;
        from now on the sweeps read the driver's 3-D deposit.  Returns the
        number of beam macro-particles."""
        self._driver_c = driver.to_c()
        n = ctypes.c_int64()
        check(self.L.qss_driver_seed(self.h, ctypes.byref(self._driver_c), ctypes.byref(n)),
              "qss_driver_seed")
        self._driver_n = n.value
        return n.value

    def MoveBeamParticles(self):
        """Mesh::MoveBeamParticles (movebeam.cpp:14-660): push through the
        stored wake and deposit rho_b, Jx_b into the 3-D node arrays."""
        check(self.L.qss_driver_push(self.h), "qss_driver_push")

    def Step(self) -> np.ndarray:
        """Domain::Step (step.cpp:23-105): MoveBeamParticles, then
        MoveAllSplitLayers; returns the sweep's residuals."""
        res = np.zeros(self.params.nslices * self.params.nx_split)
        check(self.L.qss_time_step(self.h, _p(res)), "qss_time_step")
        return res

    def driver_particles(self) -> dict:
        n = self._driver_n
        d = {"cell": np.zeros((n, 3), dtype=np.int32), "frac": np.zeros((n, 3)),
             "mom": np.zeros((n, 3)), "weight": np.zeros(n),
             "bunch": np.zeros(n, dtype=np.int32)}
        if n:
            check(self.L.qss_driver_get(self.h, d["cell"].ctypes.data_as(_i32p),
                                        _p(d["frac"]), _p(d["mom"]), _p(d["weight"]),
                                        d["bunch"].ctypes.data_as(_i32p)), "qss_driver_get")
        return d

    def driver_set(self, cell, frac, mom, weight, bunch):
        cell = np.ascontiguousarray(cell, dtype=np.int32).reshape(-1, 3)
        frac = np.ascontiguousarray(frac, dtype=np.float64).reshape(-1, 3)
        mom = np.ascontiguousarray(mom, dtype=np.float64).reshape(-1, 3)
        weight = np.ascontiguousarray(weight, dtype=np.float64)
        bunch = np.ascontiguousarray(bunch, dtype=np.int32)
        check(self.L.qss_driver_set(self.h, len(weight), cell.ctypes.data_as(_i32p), _p(frac),
                                    _p(mom), _p(weight), bunch.ctypes.data_as(_i32p)),
              "qss_driver_set")
        self._driver_n = len(weight)

    def driver_density(self, layer: int):
        rho, jx = self._plane(), self._plane()
        check(self.L.qss_driver_density(self.h, int(layer), _p(rho), _p(jx)),
              "qss_driver_density")
        return rho, jx

    def driver_timing(self) -> dict:
        c, p = ctypes.c_double(), ctypes.c_double()
        check(self.L.qss_driver_timing(self.h, ctypes.byref(c), ctypes.byref(p)),
              "qss_driver_timing")
        return {"ms_clear": c.value, "ms_push": p.value}

    def history_ptr(self, name: str) -> int:
        p = ctypes.c_void_p()
        check(self.L.qss_history_ptr(self.h, FIELD_ID[name], ctypes.byref(p)),
              "qss_history_ptr")
        return p.value

    def particles(self) -> dict:
        n = self.n_particles
        d = {k: np.zeros(n) for k in ("fy", "fz", "px", "py", "pz")}
        d["cj"] = np.zeros(n, dtype=np.int32)
        d["ck"] = np.zeros(n, dtype=np.int32)
        d["id"] = np.zeros(n, dtype=np.int64)
        check(self.L.qss_get_particles(self.h, d["cj"].ctypes.data_as(_i32p),
                                       d["ck"].ctypes.data_as(_i32p), _p(d["fy"]),
                                       _p(d["fz"]), _p(d["px"]), _p(d["py"]), _p(d["pz"]),
                                       d["id"].ctypes.data_as(_i64p)), "qss_get_particles")
        d["y"] = d["cj"] + d["fy"]
        d["z"] = d["ck"] + d["fz"]
        return d

    def set_particles(self, cj, ck, fy, fz, px, py, pz, pid):
        ii = [np.ascontiguousarray(a, dtype=np.int32) for a in (cj, ck)]
        ff = [np.ascontiguousarray(a, dtype=np.float64) for a in (fy, fz, px, py, pz)]
        pid = np.ascontiguousarray(pid, dtype=np.int64)
        check(self.L.qss_set_particles(self.h, len(pid), ii[0].ctypes.data_as(_i32p),
                                       ii[1].ctypes.data_as(_i32p), *[_p(a) for a in ff],
                                       pid.ctypes.data_as(_i64p)), "qss_set_particles")

    def layer_ptr(self, name: str, layer: str = "C") -> int:
        p = ctypes.c_void_p()
        check(self.L.qss_layer_ptr(self.h, 0 if layer == "C" else 1, FIELD_ID[name],
                                   ctypes.byref(p)), "qss_layer_ptr")
        return p.value

    def transform2d(self, x: np.ndarray, kind_z: int, kind_y: int,
                    scale: float = 1.0) -> np.ndarray:
        """scale * T_kz . x . T_ky on the GPU (CUDA_WRAP_fourierHalfInteger2D,
        half_integer2D.cu:583) of a host (nz, ny) plane.  The C-layer Rho and
        Psi planes serve as device staging (their contents are restored)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.shape != (self.nz, self.ny):
            raise QssError(-1, "transform2d", "plane shape must be (nz, ny)")
        saved = (self.get_field("Rho"), self.get_field("Psi"))
        self.set_field("Rho", x)
        check(self.L.qss_transform2d(self.h, int(kind_z), int(kind_y), self.layer_ptr("Rho"),
                                     self.layer_ptr("Psi"), float(scale)), "qss_transform2d")
        out = self.get_field("Psi")
        self.set_field("Rho", saved[0])
        self.set_field("Psi", saved[1])
        return out

    # -- transform tile shapes --------------------------------------------
    def autotune(self, reps: int = 3) -> dict:
        """Time every compiled transform tile shape for every launch shape
        and keep the fastest (FFTW_MEASURE-style planning)."""
        nv = self.gemm_variants(self.params.precision)
        tab = (ctypes.c_float * (2 * 8 * nv))()
        check(self.L.qss_autotune(self.h, int(reps), tab), "qss_autotune")
        out = self.gemm_choice()
        out["table_us"] = [[[round(tab[(p * 8 + b) * nv + v] * 1e3, 1) for v in range(nv)]
                            for b in range(8)] for p in range(2)]
        return out

    def gemm_choice(self) -> dict:
        arr = (ctypes.c_int * 18)()
        check(self.L.qss_gemm_choice(self.h, arr), "qss_gemm_choice")
        out = {"y_pass": list(arr[1:9]), "z_pass": list(arr[10:18]),
               "particle_split": self.particle_split,
               "transform_method": self.transform_method}
        fa = (ctypes.c_int * 36)()
        check(self.L.qss_fact_choice(self.h, fa), "qss_fact_choice")
        out["fact_y_pass"] = list(fa[1:9])
        out["fact_z_pass"] = list(fa[10:18])
        out["fused_y_pass"] = list(fa[19:27])
        out["fused_z_pass"] = list(fa[28:36])
        return out

    def set_gemm_choice(self, choice: dict) -> None:
        """Install a tuning choice (as returned by gemm_choice/autotune):
        transform tile shapes and, when present, the push work-groups per tile."""
        arr = (ctypes.c_int * 18)()
        for i, v in enumerate(choice["y_pass"]):
            arr[1 + i] = v
        for i, v in enumerate(choice["z_pass"]):
            arr[10 + i] = v
        check(self.L.qss_set_gemm_choice(self.h, arr), "qss_set_gemm_choice")
        if "fact_y_pass" in choice:
            fa = (ctypes.c_int * 36)()
            cur = (ctypes.c_int * 36)()
            check(self.L.qss_fact_choice(self.h, cur), "qss_fact_choice")
            for i in range(36):
                fa[i] = cur[i]
            for blk, key in enumerate(("fact_y_pass", "fact_z_pass", "fused_y_pass",
                                       "fused_z_pass")):
                for i, v in enumerate(choice.get(key, ())):
                    fa[9 * blk + 1 + i] = v
            check(self.L.qss_set_fact_choice(self.h, fa), "qss_set_fact_choice")
        if "particle_split" in choice:
            self.particle_split = choice["particle_split"]

    @property
    def particle_split(self) -> int:
        """Work-groups sharing one 16x16-cell tile in the tiled push."""
        return int(self.L.qss_particle_split(self.h))

    @particle_split.setter
    def particle_split(self, n: int) -> None:
        check(self.L.qss_set_particle_split(self.h, int(n)), "qss_set_particle_split")

    def set_gemm_variant(self, v: int):
        check(self.L.qss_set_gemm_variant(self.h, int(v)), "qss_set_gemm_variant")

    @property
    def transform_method(self) -> int:
        """QSS_TRANSFORM_DENSE (0) or QSS_TRANSFORM_FACTORED (1)."""
        return self.L.qss_transform_method(self.h)

    @transform_method.setter
    def transform_method(self, method: int) -> None:
        check(self.L.qss_set_transform_method(self.h, int(method)), "qss_set_transform_method")

    @staticmethod
    def gemm_variants(precision: int = 0) -> int:
        """Number of transform tile shapes compiled for a precision mode."""
        n = _lib.load().qss_gemm_variants(int(precision))
        if n < 0:
            raise QssError(n, "qss_gemm_variants", "unknown precision")
        return n

    # -- differentiation ---------------------------------------------------
    def set_bins(self, bins) -> None:
        """Replace the driver's piecewise-constant longitudinal profile."""
        b = np.ascontiguousarray(bins, dtype=np.float64)
        check(self.L.qss_set_bins(self.h, _p(b), int(b.size)), "qss_set_bins")
        self.checkpoint_generation += 1

    def sweep_checkpointed(self, nrun: int, every: int = 32) -> np.ndarray:
        """Sweep layers [0, nrun) from a fresh front, keeping a checkpoint every
        `every` layers for sweep_vjp (every = 0: recording every layer's tape
        instead, so that sweep_vjp recomputes nothing; ~25 MB per layer at
        256^2 x 4 ppc).  Returns the layer residuals."""
        res = np.zeros(nrun)
        check(self.L.qss_sweep_checkpointed(self.h, int(nrun), int(every), _p(res)),
              "qss_sweep_checkpointed")
        self.checkpoint_generation += 1
        return res

    def sweep_vjp(self, ex_bar=None, device_ptr: Optional[int] = None) -> np.ndarray:
        """dLoss/dbins of the preceding checkpointed sweep, given dLoss/dEx of
        its layers (host array (nrun, nz, ny), or a device pointer)."""
        g = np.zeros(len(self.params.bins) if self.params.bins is not None else 0)
        if device_ptr is not None:
            rc = self.L.qss_sweep_vjp(self.h, ctypes.c_void_p(device_ptr), 0, _p(g))
        elif ex_bar is not None:
            e = np.ascontiguousarray(ex_bar, dtype=np.float64)
            rc = self.L.qss_sweep_vjp(self.h, e.ctypes.data_as(ctypes.c_void_p), 1, _p(g))
        else:
            rc = self.L.qss_sweep_vjp(self.h, None, 0, _p(g))
        check(rc, "qss_sweep_vjp")
        return g

    def copy_history(self, name: str, first: int, count: int, dst_ptr: int) -> None:
        """Device-to-device copy of history planes into memory at dst_ptr."""
        check(self.L.qss_copy_history(self.h, FIELD_ID[name], int(first), int(count),
                                      ctypes.c_void_p(dst_ptr)), "qss_copy_history")

    @staticmethod
    def device_memory(device: int = 0) -> tuple:
        """(used, total) bytes of device memory on `device`."""
        u, t = ctypes.c_int64(), ctypes.c_int64()
        check(_lib.load().qss_device_memory(int(device), ctypes.byref(u), ctypes.byref(t)),
              "qss_device_memory")
        return u.value, t.value

    # -- timing ------------------------------------------------------------
    def enable_timing(self, on: bool = True):
        check(self.L.qss_enable_timing(self.h, int(on)), "qss_enable_timing")

    def timing(self) -> dict:
        tot = ctypes.c_double()
        ms = (ctypes.c_double * 6)()
        n = (ctypes.c_int64 * 6)()
        fl = ctypes.c_double()
        check(self.L.qss_timing(self.h, ctypes.byref(tot), ms, n, ctypes.byref(fl)),
              "qss_timing")
        names = ["push", "gemm", "spectral", "beam", "sort", "reduce"]
        return {"ms_total": tot.value, "gemm_flops": fl.value,
                "ms": {k: ms[i] for i, k in enumerate(names)},
                "launches": {k: n[i] for i, k in enumerate(names)}}
