#!/usr/bin/env python3
"""Benchmark of the quasi-static xi-slice sweep on one MI355X.

A "step" is one layer step (one xi slice: Guess, nIter x (push+deposit,
Iterate), full-step push+deposit, final Iterate) of the 512x512-cell,
4-particles-per-cell workload (1,048,576 plasma electrons per slice, 2048
slices of dxi = 0.01, driver + witness bunches).  The sweep starts at the
front; `--warmup` slices run untimed, then `--steps` slices are timed (the
defaults cover the whole 2048-slice sweep).  Everything is resident in HBM
before the timed region: plasma particles are seeded on the device and the
beam is deposited per slice on the device.

value = particle-slice pushes/s of the plain FP64 sweep = plasma particles
per slice x timed slices / timed seconds (one push = one particle advanced
across one slice with all nIter+1 = 3 passes), layer-step pairs replayed
from a captured hipGraph; the same window is also timed with plain launches
and with a HIP event pair around every launch (the per-kernel times of the
roofline come from that instrumented window).  The same timed window is run
in each mixed-precision mode ("mixed": FP32 push with FP32-MFMA or
split-BF16-MFMA transforms) and each mode's whole sweep is compared with the
FP64 sweep: relative L2 over every 8th layer of Ex (the longitudinal field,
E_z in the usual wakefield notation) and psi.

roofline: the dominant kernel is the transform GEMM; achieved = the flops of
the transform method in use (2*M*N*K per GEMM pass and field: 4N^3 per 2-D
transform with the dense tables, 2N^3 with the factored form, whose radix-2
step halves the multiply-adds; 35 2-D transforms per slice) / its summed
duration over the instrumented window, each launch timed by a HIP event pair
that the dispatch itself stamps with its start and end (hipExtLaunchKernel,
on the context's stream), against MI355X's 78.6 TFLOP/s FP64 MFMA peak;
dense_equivalent_tflops restates the rate in the dense tables' flops.
traffic = HBM bytes per launch from the committed rocprofv3
FETCH_SIZE/WRITE_SIZE summary of that kernel.  roofline_push: the push+deposit
kernel against the 8 TB/s HBM peak, its algorithmic bytes per launch being
Np*52 B of particle state + 16 N^2 planes of field reads / deposit writes.

driver: the evolving driver (3-D beam push between sweeps) on the same
workload: seeding, one time step (push + sweep) and the push kernel's time
per push with its HBM roofline (see driver_run).

cpu_baseline: the FP64 oracle (FFT-form DCT-IV, OpenMP) on the host cores
(cpu_baseline_1core: the same on one core, a shorter sample),
on the first slices of the same workload, reported in the same unit.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"fp64": 78.6, "mixed32": 157.3, "bf16x3": 2500.0, "bf16x6": 2500.0}
MFMA_KIND = {"fp64": "FP64 MFMA", "mixed32": "FP32 MFMA", "bf16x3": "BF16 MFMA (3 products)",
             "bf16x6": "BF16 MFMA (6 products)"}
HW_FLOP_FACTOR = {"fp64": 1, "mixed32": 1, "bf16x3": 3, "bf16x6": 6}
WORKLOAD = ("512x512 cells over 20x20 k_p^-1, 4 ppc (1,048,576 plasma electrons/slice), "
            "dxi=0.01, 2048 slices, driver nb=4 sr=0.25 sxi=0.7 + witness nb=8 sr=0.1 "
            "sxi=0.1, CIC, nIter=2")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2028)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--lead-in", type=int, default=-1,
                    help="untimed slices swept before --warmup, so that a short timed window "
                         "sits inside the wake (default: up to the driver centre, slice 205, "
                         "as far as --warmup + --steps leave room)")
    ap.add_argument("--cpu-slices", type=int, default=96,
                    help="slices of the CPU baseline sample (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="oracle threads (default: the cores this process may run on, at "
                         "most OMP_NUM_THREADS)")
    ap.add_argument("--no-small-configs", action="store_true",
                    help="skip the full 64^2 and 256^2 sweeps on the GPU and the CPU")
    ap.add_argument("--mixed", default="mixed32,bf16x6,bf16x3",
                    help="comma list of mixed modes to time and compare ('' = none)")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the whole-sweep accuracy comparison of the mixed modes")
    ap.add_argument("--no-timing-events", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="time plain launches only (value from the eager window, adjoint leg "
                         "without graph replay)")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the adjoint (gradient) timing on the 256^2 workload")
    ap.add_argument("--fd-bins", type=int, default=16,
                    help="bins checked against central differences in the adjoint section")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched parameter scan (64 wakes on 4 grid sizes)")
    ap.add_argument("--batch-slices", type=int, default=512)
    ap.add_argument("--driver-pushes", type=int, default=10,
                    help="evolving-driver pushes timed after one time step (0: skip)")
    ap.add_argument("--gemm-tiles", default="", metavar="Y/Z[/FY/FZ]",
                    help="FP64 transform tile variants per batch size 1..8 for the y and z "
                         "passes, e.g. 11,12,11,13,11,14,0,0/11,12,12,13,11,15,0,2, then "
                         "optionally those of the fused factored kernels "
                         "(default: the autotuner's choice)")
    ap.add_argument("--transform", default="auto", choices=("auto", "dense", "factored"),
                    help="2-D transform method of the FP64 sweeps (auto: the context's "
                         "default, factored on grids of at least 256^2)")
    ap.add_argument("--particle-split", type=int, default=0, choices=(0, 1, 2, 4),
                    help="work-groups per tile of the push (0: the autotuner's choice)")
    ap.add_argument("--library-gemm", action="store_true",
                    help="only time the platform FP64 GEMM on the transform shapes")
    ap.add_argument("--driver-only", action="store_true",
                    help="only the evolving-driver pushes, through an empty wake (no sweep: "
                         "a short run for hardware counters; the push reads the same bytes)")
    ap.add_argument("--batch-wakes", type=int, default=16, help="wakes per grid size")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 5 times (a kernel of known "
                         "HBM traffic that scales the FETCH_SIZE/WRITE_SIZE counters)")
    return ap.parse_args()


def copy_calibration(mib: int) -> int:
    """Plain device copies of `mib` MiB (read mib, write mib): the reference
    kernel for turning rocprof's FETCH_SIZE / WRITE_SIZE counters into bytes."""
    import torch
    n = mib * (1 << 20) // 8
    x = torch.arange(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_mib": mib, "launches": 5,
                      "ok": bool(torch.equal(x[-4:], y[-4:]))}), flush=True)
    return 0


def profiled_traffic(name):
    """HBM bytes per launch of a kernel from its committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, scaled by a plain copy's counters);
    None when the summary is absent."""
    import re
    path = os.path.join(ROOT, "profiles", name + "_pmc.txt")
    try:
        with open(path) as f:
            m = re.search(r"^traffic_bytes_per_launch\s*=\s*([0-9.eE+]+)", f.read(), re.M)
    except OSError:
        return None
    return float(m.group(1)) if m else None


def host_cpu():
    """CPU model, logical CPUs of the host, and the cores this process may run
    on (its affinity mask: the share of the host a job is given)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return {"model": model, "nproc": os.cpu_count(), "cores_available": avail}


def cpu_threads(a):
    """All the cores this process may use, capped by OMP_NUM_THREADS when the
    job sets it (the share of a shared host that the job is given)."""
    if a.cpu_threads:
        return a.cpu_threads
    n = host_cpu()["cores_available"]
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def cpu_baseline(params, nslices, nthreads, what=None):
    from oracle import oracle as O
    O.build()
    o = O.Oracle(params, use_fft=True, nthreads=nthreads, history=())
    t0 = time.perf_counter()
    st, _ = o.sweep_range(0, nslices)
    dt = time.perf_counter() - t0
    n = o.n_particles
    o.close()
    cpu = host_cpu()
    return {"value": n * nslices / dt, "unit": "particle-slice pushes/s", "cores": nthreads,
            "kind": "port", "slices_per_s": nslices / dt, "seconds": dt,
            "cpu_model": cpu["model"], "nproc": cpu["nproc"],
            "cores_available": cpu["cores_available"],
            "sample": (what or "first %d slices behind the front of the same 512x512x4ppc "
                       "workload" % nslices) +
                      " (FP64 oracle, FFT-form DCT-IV, %d OpenMP threads, %.1f s)"
                      % (nthreads, dt)}


def small_configs(a):
    """The other two single-wake FP64 workloads end to end: the 64^2 linear
    wake (200 slices) and the 256^2 blowout (1024 slices), each swept whole on
    the GPU (graph replay, after one untimed sweep) and by the oracle on the
    host cores."""
    from tertia_amd.mesh import Mesh
    from tertia_amd.params import blowout_256, linear_wake_64
    out = {}
    nthr = cpu_threads(a)
    for name, p in (("linear_wake_64", linear_wake_64()), ("blowout_256", blowout_256())):
        m = Mesh(p, history=())
        try:
            m.autotune(2)
            m.enable_graph(True)
            m.MoveAllSplitLayers()
            t0 = time.perf_counter()
            m.MoveAllSplitLayers()
            dt = time.perf_counter() - t0
            npart = m.n_particles
        finally:
            m.close()
        cpu = cpu_baseline(p, p.nslices, nthr, what="the whole %d-slice sweep" % p.nslices)
        out[name] = {"workload": "%dx%d x 4 ppc (%d electrons/slice), %d slices, FP64"
                                 % (p.ny, p.nz, npart, p.nslices),
                     "gpu_pushes_per_s": npart * p.nslices / dt,
                     "gpu_slices_per_s": p.nslices / dt, "gpu_s": dt,
                     "cpu": cpu, "gpu_over_cpu": cpu["seconds"] / dt}
    return out


def timed_run(mode, a):
    """Autotune, then time `steps` slices after `warmup` slices three ways on
    the same context: (1) launches bracketed by HIP events (per-kernel times
    for the roofline), (2) plain launches, (3) layer-step pairs replayed from
    a captured hipGraph (the production path; `value` is measured here)."""
    from tertia_amd.params import PRECISIONS, two_bunch_512
    from tertia_amd.mesh import Mesh
    p = two_bunch_512(precision=PRECISIONS[mode])
    m = Mesh(p, history=())
    out = {}
    try:
        if a.transform != "auto" and mode == "fp64":
            m.transform_method = 1 if a.transform == "factored" else 0
        choice = m.autotune(3)
        if a.gemm_tiles and mode == "fp64":
            # a fixed tile table (y-pass and z-pass variant per batch size 1..8;
            # then optionally those of the fused factored kernels) instead of
            # the tuned one, e.g. to repeat a run's kernels exactly
            parts = [[int(v) for v in part.split(",")] for part in a.gemm_tiles.split("/")]
            c = m.gemm_choice()
            c.update({"y_pass": parts[0], "z_pass": parts[1]})
            if len(parts) == 4:
                c.update({"fused_y_pass": parts[2], "fused_z_pass": parts[3]})
            m.set_gemm_choice(c)
            choice.update(m.gemm_choice())
        if a.particle_split:
            # work-groups per push tile fixed instead of tuned (the tuned
            # choice can flip between runs where two splits time alike)
            m.particle_split = a.particle_split
            choice["particle_split"] = m.particle_split
        npart = m.n_particles

        def window(events, graph):
            m.enable_graph(graph)
            m.reset()
            first = a.lead_in + a.warmup
            if first:
                m.sweep_range(0, first)
            m.enable_timing(events)
            t0 = time.perf_counter()
            res = m.sweep_range(first, a.steps)   # synchronises at the end
            dt = time.perf_counter() - t0
            tm = m.timing() if events else None
            m.enable_timing(False)
            return dt, res, tm

        dt_ev, _, tm = window(True, False) if not a.no_timing_events else (None, None, None)
        dt_plain, res_plain, _ = window(False, False)
        if a.no_graph:
            dt, res = dt_plain, res_plain
        else:
            dt, res, _ = window(False, True)
    finally:
        m.close()
    out.update({"pushes_per_s": npart * a.steps / dt, "slices_per_s": a.steps / dt,
                "ms_per_step": dt * 1e3 / a.steps,
                "ms_per_step_eager": dt_plain * 1e3 / a.steps,
                "max_residual": float(max(res)), "gemm_tiles": choice})
    if tm is not None:
        gemm_ms = tm["ms"]["gemm"]
        achieved = tm["gemm_flops"] / (gemm_ms * 1e-3) / 1e12 if gemm_ms > 0 else 0.0
        out["ms_per_step_events"] = dt_ev * 1e3 / a.steps
        out["gemm"] = {"achieved_tflops": achieved, "avg_launch_us":
                       gemm_ms * 1e3 / max(1, tm["launches"]["gemm"]),
                       "launches": tm["launches"]["gemm"], "flops": tm["gemm_flops"],
                       "ms": gemm_ms}
        # push kernel: algorithmic bytes per pass = Np * 52 B (y, z, px, py, pz
        # + cell id) + N^2 * (12 field planes read + 4 deposit planes written)
        fb = 4 if mode != "fp64" else 8
        nbytes = npart * 52 + p.ny * p.nz * 16 * fb
        push_ms = tm["ms"]["push"]
        nl = max(1, tm["launches"]["push"])
        out["push"] = {"avg_launch_us": push_ms * 1e3 / nl, "launches": nl,
                       "bytes_per_launch": nbytes,
                       "achieved_gbs": nbytes * nl / (push_ms * 1e-3) / 1e9 if push_ms > 0 else 0.0}
        out["kernel_ms"] = tm["ms"]
        out["kernel_launches"] = tm["launches"]
        out["event_window_ms"] = tm["ms_total"]
    return out


ACC_FIELDS = ("Ex", "Ey", "Ez", "Bx", "By", "Bz", "Psi")


def sampled_history(mode, every):
    """Whole sweep (from the front) storing Ex..Bz and psi; every `every`-th layer."""
    from tertia_amd.params import PRECISIONS, two_bunch_512
    from tertia_amd.mesh import Mesh
    p = two_bunch_512(precision=PRECISIONS[mode])
    m = Mesh(p, history=ACC_FIELDS)
    try:
        m.MoveAllSplitLayers()
        idx = range(every - 1, p.nslices, every)
        return {f: np.stack([m.history(f, s) for s in idx]) for f in ACC_FIELDS}
    finally:
        m.close()


def adjoint_run(fd_bins: int, graph: bool = True):
    """256x256 x 4 ppc (262,144 electrons/slice), FP64, 256 profile bins over
    xi in [0,4]; loss = -mean Ex + 1e-3 var Ex over xi in [6,6.5], r < 0.2,
    which only depends on the first 650 layers.  Times the plain forward
    sweep of those layers and the gradient (checkpointed forward, every 32
    layers, + reverse sweep), all with graph replay, and checks the gradient
    against central differences (h = 1e-6) on `fd_bins` random bins."""
    from tertia_amd.loss import witness_loss_np
    from tertia_amd.mesh import Mesh
    from tertia_amd.params import profile_gradient_256
    p = profile_gradient_256()
    nrun, every = 650, 32
    win = dict(xi0=6.0, xi1=6.5, rmax=0.2)
    m = Mesh(p, history=("Ex",))
    try:
        m.autotune(2)
        # layer-step pairs (forward sweeps) and checkpoint segments (reverse)
        # replayed from hipGraphs (--no-graph: launched eagerly)
        m.enable_graph(graph)
        used0, _ = Mesh.device_memory()

        def ex_hist():
            return np.stack([m.history("Ex", s) for s in range(nrun)])

        m.reset()
        m.sweep_range(0, 5)                      # warm up
        m.reset()
        t0 = time.perf_counter()
        m.sweep_range(0, nrun)
        t_fwd = time.perf_counter() - t0
        loss, _ = witness_loss_np(ex_hist(), p, **win)
        t0 = time.perf_counter()
        m.sweep_checkpointed(nrun, every)
        t_ck = time.perf_counter() - t0
        loss2, gex = witness_loss_np(ex_hist(), p, **win)
        t0 = time.perf_counter()
        g = m.sweep_vjp(gex)
        t_vjp = time.perf_counter() - t0
        # every layer's tape recorded by the forward (HBM: ~17 GB): the
        # reverse recomputes nothing
        t0 = time.perf_counter()
        m.sweep_checkpointed(nrun, 0)
        t_tf = time.perf_counter() - t0
        t0 = time.perf_counter()
        g_taped = m.sweep_vjp(gex)
        t_tr = time.perf_counter() - t0
        used_taped, _ = Mesh.device_memory()
        # the same reverse sweep with a HIP event pair per launch: breakdown
        m.sweep_checkpointed(nrun, every)
        m.enable_timing(True)
        m.sweep_vjp(gex)
        tm = m.timing()
        m.enable_timing(False)
        used1, _ = Mesh.device_memory()
        fd = None
        if fd_bins > 0:
            rng = np.random.default_rng(2026)
            idx = sorted(rng.choice(np.arange(len(p.bins)), fd_bins, replace=False).tolist())
            base = np.array(p.bins)
            gf = []
            for i in idx:
                h = 1e-6 * max(1.0, abs(base[i]))
                vals = []
                for sgn in (1, -1):
                    b = base.copy()
                    b[i] += sgn * h
                    m.set_bins(b)
                    m.reset()
                    m.sweep_range(0, nrun)
                    vals.append(witness_loss_np(ex_hist(), p, **win)[0])
                gf.append((vals[0] - vals[1]) / (2 * h))
            m.set_bins(base)
            gf = np.array(gf)
            scale = float(np.abs(g).max())
            fd = {"bins": idx, "max_rel_err": float(np.abs(g[idx] - gf).max() / scale),
                  "grad_max": scale, "h": 1e-6}
    finally:
        m.close()
    return {"workload": "256x256 over 16x16 k_p^-1, 4 ppc (262,144 electrons/slice), dxi=0.01,"
                        " FP64, 256 profile bins on xi in [0,4], loss on xi in [6,6.5] r<0.2 "
                        "(first 650 layers)",
            "mode": "taped: the forward records every layer's tape in HBM (every = 0), "
                    "the reverse recomputes nothing",
            "forward_s": t_fwd, "taped_forward_s": t_tf, "reverse_s": t_tr,
            "adjoint_over_forward": t_tr / t_fwd,
            "gradient_over_forward": (t_tf + t_tr) / t_fwd,
            "device_bytes_in_use_taped": used_taped,
            "checkpointed": {"every": every, "forward_s": t_ck, "reverse_s": t_vjp,
                             "adjoint_over_forward": t_vjp / t_fwd,
                             "gradient_over_forward": (t_ck + t_vjp) / t_fwd,
                             "gradient_equals_taped": bool(np.array_equal(g_taped, g))},
            "note": "adjoint = the reverse sweep; gradient = the recording forward + the "
                    "reverse sweep, both against the plain forward; checkpointed: state "
                    "every 32 layers, each segment recomputed in the reverse; forward sweeps "
                    "by graph-replayed layer pairs, checkpointed reverse segments by one "
                    "graph each; reverse_kernel_ms from a further, event-timed (eager) "
                    "checkpointed reverse sweep",
            "reverse_kernel_ms": tm["ms"], "reverse_launches": tm["launches"],
            "checkpoint_every": every,
            "loss": loss, "loss_checkpointed": loss2,
            "device_bytes_in_use": used1, "device_bytes_before": used0,
            "fd_check": fd}


def batch_run(nslices: int, per_size: int):
    """Parameter scan: per_size wakes on each of 128^2, 192^2, 256^2, 384^2
    (4 ppc, nslices layers, random driver parameters).  Batched: one context
    per grid size advancing all its wakes with every launch; sequential: the
    same wakes one context at a time (same tile-shape tuning per size)."""
    from tertia_amd.mesh import Mesh
    from tertia_amd.params import Params, parameter_scan
    groups = parameter_scan(per_size=per_size, nslices=nslices)
    out = {"workload": "%d wakes: %d each on 128^2, 192^2, 256^2, 384^2 (h = 1/16), 4 ppc, "
                       "%d layers of dxi = 0.02, FP64; sigma_r~U[0.2,0.8], "
                       "sigma_xi~U[0.5,1.5], nb~U[0.5,8]" % (4 * per_size, per_size, nslices),
           "groups": []}
    t_b = t_s = 0.0
    for p in groups:
        m = Mesh(p, history=())
        try:
            m.autotune(2)
            m.reset()
            m.sweep_range(0, 2)
            m.reset()
            t0 = time.perf_counter()
            m.sweep_range(0, nslices)
            tbe = time.perf_counter() - t0
            m.enable_graph(True)          # the slice step replayed from a hipGraph
            m.reset()
            t0 = time.perf_counter()
            m.sweep_range(0, nslices)
            tb = time.perf_counter() - t0
        finally:
            m.close()
        ts = 0.0
        choice = None
        for w in p.wakes:
            ps = Params(ny=p.ny, nz=p.nz, ly=p.ly, lz=p.lz, hx=p.hx, nslices=nslices,
                        bunches=list(w))
            m = Mesh(ps, history=())
            try:
                if choice is None:
                    choice = m.autotune(2)
                else:
                    m.set_gemm_choice(choice)
                m.reset()
                t0 = time.perf_counter()
                m.sweep_range(0, nslices)
                ts += time.perf_counter() - t0
            finally:
                m.close()
        out["groups"].append({"n": p.ny, "wakes": len(p.wakes), "batched_s": tb,
                              "batched_eager_s": tbe, "sequential_s": ts})
        t_b += tb
        t_s += ts
    total = 4 * per_size * nslices
    out.update({"batched_slices_per_s": total / t_b, "sequential_slices_per_s": total / t_s,
                "speedup": t_s / t_b})
    # all four grid sizes at once: one batched context per size, each on its
    # own stream, swept from its own host thread (the library releases the
    # interpreter during a sweep), so the sizes' launches overlap on the GPU
    import threading
    ms = []
    try:
        for p in groups:
            m = Mesh(p, history=())
            ms.append(m)
            m.autotune(2)
            m.enable_graph(True)
            m.reset()
            m.sweep_range(0, 3)          # captures the graph before the timed threads
            m.reset()
        errs = []

        def run(m):
            try:
                m.sweep_range(0, nslices)
            except Exception as e:   # reported after the join
                errs.append(e)
        th = [threading.Thread(target=run, args=(m,)) for m in ms]
        t0 = time.perf_counter()
        for t in th:
            t.start()
        for t in th:
            t.join()
        tc = time.perf_counter() - t0
        if errs:
            raise errs[0]
    finally:
        for m in ms:
            m.close()
    out.update({"concurrent_s": tc, "concurrent_slices_per_s": total / tc,
                "speedup_concurrent": t_s / tc})
    return out


def library_gemm(reps: int = 50):
    """The platform's FP64 GEMM (torch.matmul -> rocBLAS/hipBLASLt) on the
    transform's shapes: b fields of 512x512 times a 512x512 matrix, as the
    y pass (X.T, M = 512 b) and the z pass (T.X batched), timed with events
    on torch's stream.  A reference point for the transform kernel's
    fraction of peak, not part of the sweep."""
    import torch
    torch.manual_seed(0)
    dev = "cuda:0"
    T = torch.randn(512, 512, dtype=torch.float64, device=dev)
    out = {}
    for b in (1, 2, 3, 6, 8):
        X = torch.randn(b * 512, 512, dtype=torch.float64, device=dev)
        Xb = X.view(b, 512, 512)
        for name, fn in (("y_pass", lambda: X @ T), ("z_pass", lambda: torch.matmul(T, Xb))):
            for _ in range(5):
                fn()
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                fn()
            e1.record()
            torch.cuda.synchronize()
            us = e0.elapsed_time(e1) / reps * 1e3
            out["%s_b%d" % (name, b)] = {"us": us, "tflops": 2.0 * b * 512 ** 3 / us / 1e6}
    return out


def driver_run(pushes: int, sweep: bool = True):
    """Evolving driver on the 512^2 workload: the driver and witness seeded
    as 3-D beam macro-particles (4 per cell of their support, px0 = 2000,
    electrons, Ts = 2, radiation reaction above gamma = 100), one time step
    (push through the empty wake + a full FP64 sweep that stores Ex..Bz of
    all 2048 layers), then `pushes` pushes through that wake.  The push
    kernel is HBM-bound; its algorithmic bytes are the particle state it
    reads (cell 3 x int32, fractions and momenta 6 x f64, weight f64, bunch
    int32: 72 B) and writes (60 B) per particle."""
    from tertia_amd.mesh import Mesh
    from tertia_amd.params import Driver, two_bunch_512
    p = two_bunch_512()
    m = Mesh(p, history=("Ex", "Ey", "Ez", "Bx", "By", "Bz"))
    try:
        drv = Driver(ts=2.0, p0=[(2000.0, 0.0, 0.0)] * len(p.bunches),
                     q2m=[-1.0] * len(p.bunches))
        t0 = time.perf_counter()
        n = m.SeedBeamParticles(drv)
        t_seed = time.perf_counter() - t0
        t0 = time.perf_counter()
        if sweep:
            m.Step()
        t_step = time.perf_counter() - t0 if sweep else None
        push, clear = [], []
        for _ in range(pushes):
            m.MoveBeamParticles()
            t = m.driver_timing()
            push.append(t["ms_push"])
            clear.append(t["ms_clear"])
        ms_push = float(np.mean(push))
        bytes_state = 132.0 * n
        gbs = bytes_state / (ms_push * 1e-3) / 1e9
        return {"particles": n, "pushes": pushes, "ms_push": ms_push,
                "ms_clear": float(np.mean(clear)), "seed_s": t_seed,
                "time_step_s": t_step, "particle_pushes_per_s": n / (ms_push * 1e-3),
                "roofline": {"bound": "hbm", "achieved": gbs, "peak": 8000.0, "unit": "GB/s",
                             "frac": gbs / 8000.0, "bytes_per_launch": bytes_state,
                             "traffic": profiled_traffic("k_driver_push_512x512"),
                             "kernel": "k_driver_push (3-D beam push + deposit)"}}
    finally:
        m.close()


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU; --gpus %d is not supported "
              "(the xi sweep of one wake is sequential; independent wakes would shard "
              "one batched context per GPU)" % a.gpus,
              file=sys.stderr)
        return 2
    if a.copy_calibration:
        return copy_calibration(a.copy_calibration)
    if a.library_gemm:
        print(json.dumps({"library_gemm": library_gemm()}))
        return 0
    if a.driver_only:
        print(json.dumps({"driver": driver_run(a.driver_pushes or 10, sweep=False)}))
        return 0
    from tertia_amd.params import two_bunch_512
    ntot = two_bunch_512().nslices
    if a.lead_in < 0:
        a.lead_in = max(0, min(205, ntot - a.warmup - a.steps))
    if a.lead_in + a.warmup + a.steps > ntot:
        print("lead-in + warmup + steps must be <= %d slices" % ntot, file=sys.stderr)
        return 2
    modes = [s for s in a.mixed.split(",") if s]

    base = timed_run("fp64", a)
    print("fp64: %.3f ms/slice" % base["ms_per_step"], file=sys.stderr, flush=True)
    mixed = {}
    for mode in modes:
        mixed[mode] = timed_run(mode, a)
        print("%s: %.3f ms/slice" % (mode, mixed[mode]["ms_per_step"]), file=sys.stderr,
              flush=True)
    if modes and not a.no_accuracy:
        ref = sampled_history("fp64", 8)
        for mode in modes:
            h = sampled_history(mode, 8)
            mixed[mode]["rel_l2_vs_fp64"] = {
                f: float(np.linalg.norm(h[f] - ref[f]) / np.linalg.norm(ref[f]))
                for f in ACC_FIELDS}
            mixed[mode]["max_err_vs_fp64"] = {
                f: float(np.abs(h[f] - ref[f]).max() / np.abs(ref[f]).max())
                for f in ACC_FIELDS}
            mixed[mode]["accuracy_note"] = (
                "whole 2048-slice sweep, every 8th layer, against the FP64 GPU sweep (which "
                "matches the FP64 oracle to rel-L2 <= 2e-12 through slice 704, "
                "tests/test_gpu_two_bunch_512.py); max error relative to the field's max")
            del h
            print("%s accuracy %s" % (mode, mixed[mode]["rel_l2_vs_fp64"]), file=sys.stderr,
                  flush=True)
    for mode, r in mixed.items():
        r["speedup_vs_fp64"] = r["slices_per_s"] / base["slices_per_s"]
        if "gemm" in r:
            r["gemm"]["peak_tflops"] = PEAK_TFLOPS[mode]
            r["gemm"]["mfma"] = MFMA_KIND[mode]
            r["gemm"]["frac_of_peak_hw"] = (r["gemm"]["achieved_tflops"] *
                                            HW_FLOP_FACTOR[mode] / PEAK_TFLOPS[mode])

    line = {
        "metric": "particle-slice pushes/s at 512^2*4ppc, FP64 vs mixed",
        "value": base["pushes_per_s"],
        "unit": "particle-slice pushes/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": base["ms_per_step"],
        "ms_per_step_eager": base["ms_per_step_eager"],
        "ms_per_step_with_events": base.get("ms_per_step_events"),
        "value_window": "eager" if a.no_graph else "graph",
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded plasma lattice + frozen bi-Gaussian driver/witness)",
        "config": {"workload": WORKLOAD, "slices_per_s": base["slices_per_s"],
                   "timed_slices": [a.lead_in + a.warmup, a.lead_in + a.warmup + a.steps],
                   "timed_xi": [round((a.lead_in + a.warmup + 1) * 0.01, 2),
                                round((a.lead_in + a.warmup + a.steps) * 0.01, 2)],
                   "window_note": "%d untimed lead-in + %d warm-up slices from the front, "
                                  "then the timed slices (the driver centre is at xi = 2.1)"
                                  % (a.lead_in, a.warmup),
                   "gemm_tiles": base["gemm_tiles"], "max_residual": base["max_residual"]},
    }
    if "gemm" in base:
        g = base["gemm"]
        factored = base["gemm_tiles"].get("transform_method") == 1
        line["roofline"] = {"bound": "mfma", "achieved": g["achieved_tflops"],
                            "peak": PEAK_TFLOPS["fp64"], "unit": "TFLOP/s",
                            "frac": g["achieved_tflops"] / PEAK_TFLOPS["fp64"],
                            "traffic": profiled_traffic("k_fact_fp64_512x512" if factored
                                                        else "k_gemm_fp64_512x512"),
                            "kernel": ("k_fact_y / k_fact_z (fused factored transform GEMMs, "
                                       "2N^3 flop per 2-D transform)" if factored else
                                       "k_gemm<MathF64,...> (transform GEMM, 4N^3 flop per "
                                       "2-D transform)"),
                            "dense_equivalent_tflops": g["achieved_tflops"] * (2 if factored
                                                                               else 1),
                            "avg_launch_us": g["avg_launch_us"], "launches": g["launches"]}
        pu = base["push"]
        line["roofline_push"] = {"bound": "hbm", "achieved": pu["achieved_gbs"], "peak": 8000.0,
                                 "unit": "GB/s", "frac": pu["achieved_gbs"] / 8000.0,
                                 "traffic": profiled_traffic("k_push_tile_fp64_512x512"),
                                 "kernel": "k_push_tile<double,double,false> (push + "
                                           "deposit of Jx, Jy, Jz)",
                                 "avg_launch_us": pu["avg_launch_us"],
                                 "launches": pu["launches"],
                                 "bytes_per_launch": pu["bytes_per_launch"]}
        line["kernel_ms"] = base["kernel_ms"]
        line["event_window_ms"] = base["event_window_ms"]
    if mixed:
        line["mixed"] = mixed
    if not a.no_batch:
        line["batch"] = batch_run(a.batch_slices, a.batch_wakes)
        print("batch speedup %.2f" % line["batch"]["speedup"], file=sys.stderr, flush=True)
    if a.driver_pushes > 0:
        line["driver"] = driver_run(a.driver_pushes)
        print("driver push %.3f ms (%d particles)" % (line["driver"]["ms_push"],
                                                    line["driver"]["particles"]),
              file=sys.stderr, flush=True)
    if not a.no_adjoint:
        line["adjoint"] = adjoint_run(a.fd_bins, graph=not a.no_graph)
        ad = line["adjoint"]
        print("adjoint/forward %.2f, gradient/forward %.2f (taped); checkpointed %.2f, %.2f"
              % (ad["adjoint_over_forward"], ad["gradient_over_forward"],
                 ad["checkpointed"]["adjoint_over_forward"],
                 ad["checkpointed"]["gradient_over_forward"]),
              file=sys.stderr, flush=True)
    if not a.no_small_configs:
        line["small_configs"] = small_configs(a)
        print("small configs %s" % {k: round(v["gpu_over_cpu"], 1)
                                    for k, v in line["small_configs"].items()},
              file=sys.stderr, flush=True)
    if a.cpu_slices > 0:
        nthr = cpu_threads(a)
        line["cpu_baseline"] = cpu_baseline(two_bunch_512(), a.cpu_slices, nthr)
        # one core, as the reference runs one rank (FFTW with at most 2 threads)
        line["cpu_baseline_1core"] = cpu_baseline(two_bunch_512(), max(1, a.cpu_slices // 12),
                                                  1)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
