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

value = particle-slice pushes/s = plasma particles per slice x timed slices /
timed seconds (one push = one particle advanced across one slice with all
nIter+1 = 3 passes).

roofline: the dominant kernel is the transform GEMM (FP64 MFMA); achieved =
its algorithmic flops (2*M*N*K per GEMM pass and field) / its summed HIP-event
time over the timed region, against MI355X's 78.6 TFLOP/s FP64 peak.

cpu_baseline: the FP64 oracle (FFT-form DCT-IV, OpenMP) on the host cores,
on the first slices of the same workload, reported in the same unit.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_MFMA_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2028)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-slices", type=int, default=6,
                    help="slices of the CPU baseline sample (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-timing-events", action="store_true")
    return ap.parse_args()


def cpu_baseline(params, nslices, nthreads):
    from oracle import oracle as O
    O.build()
    o = O.Oracle(params, use_fft=True, nthreads=nthreads, history=())
    t0 = time.perf_counter()
    st, _ = o.sweep_range(0, nslices)
    dt = time.perf_counter() - t0
    n = o.n_particles
    o.close()
    return {"value": n * nslices / dt, "unit": "particle-slice pushes/s", "cores": nthreads,
            "kind": "port",
            "sample": "first %d slices behind the front of the same 512x512x4ppc workload "
                      "(FP64 oracle, FFT-form DCT-IV, %d OpenMP threads, %.1f s)"
                      % (nslices, nthreads, dt)}


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported "
              "(the sweep is strictly sequential in xi, see DESIGN.md for sharding)" % a.gpus,
              file=sys.stderr)
        return 2
    from tertia_amd.params import config2
    from tertia_amd.mesh import Mesh

    p = config2()
    if a.warmup + a.steps > p.nslices:
        print("warmup + steps must be <= %d slices" % p.nslices, file=sys.stderr)
        return 2
    m = Mesh(p, history=())
    choice = m.autotune(3)
    npart = m.n_particles
    m.reset()
    if a.warmup:
        m.sweep_range(0, a.warmup)
    if not a.no_timing_events:
        m.enable_timing(True)
    t0 = time.perf_counter()
    res = m.sweep_range(a.warmup, a.steps)   # synchronises the stream at the end
    dt = time.perf_counter() - t0
    tm = m.timing() if not a.no_timing_events else None
    m.close()

    value = npart * a.steps / dt
    line = {
        "metric": "particle-slice pushes/s at 512^2*4ppc (FP64)",
        "value": value,
        "unit": "particle-slice pushes/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dt * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded plasma lattice + frozen bi-Gaussian driver/witness)",
        "config": {"workload": "512x512 cells over 20x20 k_p^-1, 4 ppc (1,048,576 plasma "
                               "electrons/slice), dxi=0.01, 2048 slices, driver nb=4 "
                               "sr=0.25 sxi=0.7 + witness nb=8 sr=0.1 sxi=0.1, CIC, nIter=2",
                   "slices_per_s": a.steps / dt,
                   "gemm_tiles": choice,
                   "max_residual": float(max(res)) if len(res) else None},
    }
    if tm is not None:
        gemm_ms = tm["ms"]["gemm"]
        achieved = tm["gemm_flops"] / (gemm_ms * 1e-3) / 1e12 if gemm_ms > 0 else 0.0
        line["roofline"] = {"bound": "mfma", "achieved": achieved,
                            "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                            "frac": achieved / FP64_MFMA_PEAK_TFLOPS, "traffic": None,
                            "kernel": "k_gemm_f64 (transform GEMM)",
                            "avg_launch_us": gemm_ms * 1e3 / max(1, tm["launches"]["gemm"]),
                            "launches": tm["launches"]["gemm"]}
        line["kernel_ms"] = tm["ms"]
        line["event_window_ms"] = tm["ms_total"]
    if a.cpu_slices > 0:
        nthr = a.cpu_threads or min(16, os.cpu_count() or 1)
        line["cpu_baseline"] = cpu_baseline(p, a.cpu_slices, nthr)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
