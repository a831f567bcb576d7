"""Benchmark: the resident FP64 LES step (advection + Smagorinsky diffusion) on 512 x 512 x 128 cells.

One step is one call of the C-ABI entry cpp_fort_shared_les_step (F_new = F + dt * tendency, halos
included) on fields resident in HBM, submitted as one hipGraph replay; the fields alternate between two
buffer sets. Prints one JSON line:

  value      ms per LES step (lower is better), timed over --steps steps between two device syncs
  roofline   algorithmic HBM bytes of the step kernel, 8*(4P + 4I) with P padded and I interior cells,
             divided by its average duration from HIP events recorded on the context stream
  cpu_baseline  the repository's native Fortran build of the same step (oracle/_build/les_native) on the
             host cores: one warm-up step, 7 repeats of --cpu-steps steps, the median; also with 4
             threads; its warm-up step is bit-compared with the first device step
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import __graft_entry__ as graft  # noqa: E402

METRIC = "ms per LES step + HBM GB/s; mixed vs FP64 resident at 1e-10 rel err"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profiles(name: str):
    """HBM bytes per launch of `name` measured by rocprofv3 PMC runs and recorded under profiles/."""
    path = ROOT / "profiles" / "traffic.json"
    if not path.exists():
        return None
    data = json.loads(path.read_text())
    entry = data.get(name)
    return None if entry is None else entry.get("bytes_per_launch")


def poisson_section(ftc, scope, lib, fields, n=(512, 512, 128), h=(10.0, 10.0, 10.0), reps=10, check=None):
    """Pressure Poisson projection solve: the FP64 direct solve (FFT/Hartley transforms in x and y for
    periodic power-of-two sizes, FP64 z solves) vs the mixed solve (the same Hartley passes with FP32
    workspaces, two applications, where those apply; split-bf16 transforms and three applications
    otherwise; FP64 residuals and refinement), both resident, timed with HIP events over `reps` solves.
    Returns the section dict and the two solutions (interiors)."""
    pressure = graft._module("pressure")
    nx, ny, nz = n
    b = fields.divergence_rhs(nx, ny, nz, h)
    rhs = np.zeros((nx + 2, ny + 2, nz + 2), order="F")
    rhs[1:-1, 1:-1, 1:-1] = b
    out = {}
    sols = {}
    dom = scope.Domain(lib)
    hr = dom.bind(rhs, lower=(0, 0, 0))
    stream = ctypes.c_void_p(dom.stream())
    ev0, ev1 = ctypes.c_void_p(), ctypes.c_void_p()
    lib.ftc_event_create(ctypes.byref(ev0))
    lib.ftc_event_create(ctypes.byref(ev1))
    pow2 = all(d >= 4 and d & (d - 1) == 0 for d in (nx, ny))
    for name, mixed, solves in (("fp64", False, 1), ("mixed", True, 2 if pow2 else 3)):
        phi = np.zeros((nx + 4, ny + 4, nz + 4), order="F")
        hp = dom.bind(phi, lower=(-1, -1, -1))
        solver = pressure.PressureSolver(lib, dom, n, (nx * h[0], ny * h[1], nz * h[2]), mixed)
        solver.refinement(1e-12, 8, True)
        solver.solve(hp, hr)  # warm-up, uploads, and the measured residual of a converged solve
        rep = solver.report()
        solver.refinement(1e-12, solves, False)  # fixed solve count, no host synchronisation
        solver.solve(hp, hr)
        dom.wait()
        lib.ftc_event_record(ev0, stream)
        for _ in range(reps):
            solver.solve(hp, hr)
        lib.ftc_event_record(ev1, stream)
        ms = ctypes.c_float()
        lib.ftc_event_elapsed_ms(ev0, ev1, ctypes.byref(ms))
        dom.observe(hp)
        sols[name] = phi[2:-2, 2:-2, 2:-2].copy()
        solver.destroy()
        out[name] = {"ms_per_solve": round(ms.value / reps, 4), "solves": solves if mixed else 1,
                     "converged_iterations": rep.iterations, "relative_residual": rep.relative_residual}
    dom.close()
    N = nx * ny * nz
    P = (nx + 2) * (ny + 2) * (nz + 2)
    out["speedup_mixed_over_fp64"] = round(out["fp64"]["ms_per_solve"] / out["mixed"]["ms_per_solve"], 3)
    scale = np.abs(sols["fp64"]).max()
    out["max_rel_diff_mixed_vs_fp64"] = float(np.abs(sols["mixed"] - sols["fp64"]).max() / scale)
    out["fp64"]["method"] = "fft (Hartley) x/y passes + FP64 z Thomas" if pow2 else "dense FP64 MFMA transforms"
    out["mixed"]["method"] = ("Hartley passes with FP32 workspaces, FP64 residual" if pow2
                              else "split-bf16 MFMA transforms, FP64 residual")
    # The FP64 FFT solve streams the field five times (x, y forward, z, y, x inverse), each a read and a
    # write of N doubles; its floor is one read of b and one write of phi.
    pass_bytes = 5 * 2 * 8 * N
    t = out["fp64"]["ms_per_solve"] * 1e-3
    out["fp64"]["pass_bytes"] = pass_bytes
    out["fp64"]["pass_GB_per_s"] = round(pass_bytes / t / 1e9, 1)
    out["fp64"]["hbm_frac_of_pass_bytes"] = round(pass_bytes / t / 1e9 / HBM_PEAK_GBS, 4)
    out["lower_bound_bytes"] = 8 * 2 * N
    out["fp64"]["hbm_frac_of_lower_bound"] = round(out["lower_bound_bytes"] / t / 1e9 / HBM_PEAK_GBS, 4)
    out["rhs_padded_cells"] = P
    return out, sols, b


def projected_step_section(ftc, scope, lib, fields, n=(512, 512, 128), steps=10, dx=10.0, dt=0.2, cs=0.17):
    """A whole resident time step in ELMM's order (LES step, divergence into the solver's RHS, pressure
    solve, projection, the shift of phi by its reference-plane mean), all entries on the device in one
    epoch, ping-ponging two field sets: the FP64 solve against the mixed one (two or three solves, no host
    synchronisation). The context holds each step's work until the step's flush
    (FORT_SCOPE_GRAPH_QUEUE_WORK=0): the solver's launches are captured into one node, the plane sum stays
    on the device (a queued fixed-order reduction into a one-value array that the shift reads), so each
    step is one graph replay. Timed with HIP events over `steps` steps after three warm-up steps (the first
    sighting and the two ping-pong instantiations); both modes start from the same fields, and the field
    difference between them is reported after warm-up + `steps` steps."""
    pressure = graft._module("pressure")
    nx, ny, nz = n
    start = fields.les_fields(nx, ny, nz)
    out, finals = {}, {}
    sc = (ftc.dp(dx), ftc.dp(dx), ftc.dp(dx), ftc.dp(dt))
    ext = (ftc.ip(nx), ftc.ip(ny), ftc.ip(nz))
    k_ref = nz // 2 + 1  # the reference plane (1-based), at array index k_ref + 1 of phi(-1:)
    plane_lo = (ctypes.c_longlong * 3)(2, 2, k_ref + 1)
    plane_hi = (ctypes.c_longlong * 3)(nx + 2, ny + 2, k_ref + 2)
    warm = 3
    for name, mixed in (("fp64", False), ("mixed", True)):
        a = [x.copy(order="F") for x in start]
        b = [np.zeros_like(start[0], order="F") for _ in range(4)]
        rhs = np.zeros_like(start[0], order="F")
        phi = np.zeros((nx + 4, ny + 4, nz + 4), order="F")
        partial, ref = np.zeros(65536 + 256), np.zeros(1)
        held = os.environ.get("FORT_SCOPE_GRAPH_QUEUE_WORK")
        os.environ["FORT_SCOPE_GRAPH_QUEUE_WORK"] = "0"
        try:
            dom = scope.Domain(lib)
        finally:
            if held is None:
                del os.environ["FORT_SCOPE_GRAPH_QUEUE_WORK"]
            else:
                os.environ["FORT_SCOPE_GRAPH_QUEUE_WORK"] = held
        ha, hb = [dom.bind(x) for x in a], [dom.bind(x, initialized=True) for x in b]
        hr, hp = dom.bind(rhs, lower=(0, 0, 0), initialized=False), dom.bind(phi, lower=(-1, -1, -1))
        hpart, href = dom.bind(partial, initialized=False), dom.bind(ref, initialized=False)
        solver = pressure.PressureSolver(lib, dom, n, (nx * dx, ny * dx, nz * dx), mixed)
        # A fixed number of solves and no residual norms: no host synchronisation inside the step. The
        # mixed solve takes two on the Hartley passes (periodic power-of-two x and y), three otherwise.
        pow2 = all(d >= 4 and d & (d - 1) == 0 for d in (nx, ny))
        solver.refinement(1e-12, (2 if pow2 else 3) if mixed else 1, False)
        stream = ctypes.c_void_p(dom.stream())
        ev0, ev1 = ctypes.c_void_p(), ctypes.c_void_p()
        lib.ftc_event_create(ctypes.byref(ev0))
        lib.ftc_event_create(ctypes.byref(ev1))

        def step(k):
            src, dst = (ha, hb) if k % 2 == 0 else (hb, ha)
            scope.check(lib, lib.cpp_fort_shared_les_step(dom.ctx, 1, *dst, *src, *sc, ftc.dp(cs), *ext, ftc.ip(1)))
            scope.check(lib, lib.cpp_fort_shared_pressure_divergence(dom.ctx, 1, hr, *dst[:3], *sc, *ext))
            solver.solve(hp, hr)
            scope.check(lib, lib.cpp_fort_shared_pressure_projection_bc(dom.ctx, 1, *dst[:3], hp, *sc, *ext,
                                                                        ftc.ip(0), ftc.ip(0)))
            # fixpressuretoreference: phi less the mean of its reference plane, a fixed-order device sum
            # left on the device for the shift.
            scope.check(lib, lib.cpp_fort_shared_fixed_order_reduction_device(dom.ctx, 1, hp, hpart, href, ftc.ip(0),
                                                                              plane_lo, plane_hi, ftc.dp(0.0)))
            scope.check(lib, lib.cpp_fort_shared_pressure_fix_reference_sum(dom.ctx, 1, hp, href, *ext))
            dom.flush()
            return dst

        for k in range(warm):
            step(k)
        dom.wait()
        before = dom.graph_stats()
        lib.ftc_event_record(ev0, stream)
        for k in range(warm, warm + steps):
            last = step(k)
        lib.ftc_event_record(ev1, stream)
        ms = ctypes.c_float()
        lib.ftc_event_elapsed_ms(ev0, ev1, ctypes.byref(ms))
        after = dom.graph_stats()
        dom.observe(*last[:3])
        finals[name] = [x.copy() for x in ((b if (warm + steps) % 2 == 1 else a)[:3])]
        solver.destroy()
        dom.close()
        out[name] = {"ms_per_step": round(ms.value / steps, 4),
                     "graph_submissions_per_step": (after.submissions - before.submissions) / steps,
                     "graph_replays_per_step": (after.replays - before.replays) / steps}
    out["speedup_mixed_over_fp64"] = round(out["fp64"]["ms_per_step"] / out["mixed"]["ms_per_step"], 3)
    out["max_rel_field_diff_mixed_vs_fp64"] = max(
        float(np.abs(m - f).max() / np.abs(f).max()) for m, f in zip(finals["mixed"], finals["fp64"]))
    out["steps"] = warm + steps
    return out


def chain_section(ftc, scope, lib, n=4096, reps=3):
    """D = A*B*C on n x n FP64 arrays, FP64 MFMA vs split-bf16 products, resident."""
    rng = np.random.default_rng(20261123)
    A, B, C = (np.asfortranarray(rng.uniform(-1.0, 1.0, (n, n))) for _ in range(3))
    out = {}
    results = {}
    dom = scope.Domain(lib)
    ha, hb, hc = (dom.bind(X) for X in (A, B, C))
    stream = ctypes.c_void_p(dom.stream())
    ev0, ev1 = ctypes.c_void_p(), ctypes.c_void_p()
    lib.ftc_event_create(ctypes.byref(ev0))
    lib.ftc_event_create(ctypes.byref(ev1))
    for name, prec, peak in (("fp64_mfma", 0, 78.6e12), ("bf16x3", 1, 2.5e15)):
        D = np.zeros((n, n), order="F")
        hd = dom.bind(D)

        def run():
            scope.check(lib, lib.cpp_fort_shared_chain_contraction(dom.ctx, 1, hd, ha, hb, hc, ftc.ip(n),
                                                                   ftc.ip(prec)))
            dom.flush()

        run()
        dom.wait()
        lib.ftc_event_record(ev0, stream)
        for _ in range(reps):
            run()
        lib.ftc_event_record(ev1, stream)
        ms = ctypes.c_float()
        lib.ftc_event_elapsed_ms(ev0, ev1, ctypes.byref(ms))
        dom.observe(hd)
        results[name] = D
        t = ms.value / reps * 1e-3
        flops = 4.0 * n ** 3
        mfma_flops = flops * (3 if prec else 1)
        out[name] = {"ms": round(t * 1e3, 3), "tflops_fp64_equivalent": round(flops / t / 1e12, 2),
                     "mfma_utilisation": round(mfma_flops / t / peak, 4)}
    dom.close()
    return out, results, (A, B, C)


def ensemble_section(lib, fields, count=64, steps=10, tune_cache=None):
    """64 uneven subdomains (nx, ny in [32, 128], nz in [32, 96] from the seed), one projected LES step each:
    per-member plain FP64 launches vs batched FP64 vs batched mixed (autotuned), all graph-free or
    one graph per step as the mode says; plus the mixed-vs-FP64 field error after two steps."""
    ens = graft._module("ensemble")
    sizes = fields.ensemble_sizes(count)
    d, dt, cs = (10.0, 10.0, 10.0), 0.2, 0.17
    init = [fields.les_fields(*n, *d, seed=1000 + m) for m, n in enumerate(sizes)]
    out = {"members": count, "cells": int(sum(a * b * c for a, b, c in sizes))}
    finals = {}
    for name, mode in (("fp64_per_member", ens.FP64_PER_MEMBER), ("fp64_batched", ens.FP64_BATCHED),
                       ("mixed_batched", ens.MIXED_BATCHED)):
        e = ens.Ensemble(lib, sizes, d, dt, cs, mode)
        if mode != ens.FP64_PER_MEMBER:
            e.autotune(tune_cache)
        for m in range(count):
            e.upload(m, *init[m])
        e.step(2)
        e.synchronize()
        finals[name] = [e.download(m)[:3] for m in range(count)]
        stream = e.stream()
        ev0, ev1 = ctypes.c_void_p(), ctypes.c_void_p()
        lib.ftc_event_create(ctypes.byref(ev0))
        lib.ftc_event_create(ctypes.byref(ev1))
        e.step(1)
        e.synchronize()
        lib.ftc_event_record(ev0, stream)
        e.step(steps)
        lib.ftc_event_record(ev1, stream)
        ms = ctypes.c_float()
        lib.ftc_event_elapsed_ms(ev0, ev1, ctypes.byref(ms))
        info = e.info()
        e.destroy()
        out[name] = {"ms_per_step": round(ms.value / steps, 4), "launches_per_step": info.launches_per_step,
                     "les_kchunk": info.les_kchunk, "gemm_tile": info.gemm_tile,
                     "blocks_per_cu": info.blocks_per_cu, "solves": info.solves}
        out["algorithmic_bytes_per_step"] = info.algorithmic_bytes
    base = out["fp64_per_member"]["ms_per_step"]
    out["speedup_mixed_batched_over_fp64_per_member"] = round(base / out["mixed_batched"]["ms_per_step"], 3)
    out["speedup_fp64_batched_over_fp64_per_member"] = round(base / out["fp64_batched"]["ms_per_step"], 3)
    err = 0.0
    for m in range(count):
        for g, w in zip(finals["mixed_batched"][m], finals["fp64_per_member"][m]):
            err = max(err, float(np.abs(g - w).max() / np.abs(w).max()))
    out["max_rel_field_error_mixed_vs_fp64_after_2_steps"] = err
    return out


NATIVE = ROOT / "oracle" / "_build" / "les_native"


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_native(path_in, path_out, nx, ny, nz, steps, repeats, threads):
    """Runs the native Fortran step (oracle/_build/les_native) and returns its per-repeat ms per step."""
    import subprocess

    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores",
               OMP_DYNAMIC="FALSE")
    cmd = [str(NATIVE), str(nx), str(ny), str(nz), str(steps), str(repeats), str(path_in)]
    if path_out:
        cmd.append(str(path_out))
    out = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=600, check=True).stdout
    reps = [float(line.split()[-1]) for line in out.splitlines() if line.startswith("repeat ")]
    median = float(next(line.split()[-1] for line in out.splitlines() if line.startswith("median_ms_per_step")))
    return reps, median


def native_baseline(initial, first, nx, ny, nz, steps):
    """The CPU baseline: the repository's native Fortran build of the same LES step (amdflang or gfortran,
    -O3 -fopenmp -ffp-contract=off), on the same seeded fields, timed with system_clock around its step
    loop: one warm-up step, then 7 repeats of `steps` steps, the median; once with every core this process
    may use and once with 4 threads (the reference harness's default). Its warm-up step is also the
    full-size bit check of the first device step."""
    import tempfile

    stamp = NATIVE.with_name("les_native.compiler")
    compiler = stamp.read_text().strip() if stamp.exists() else "unknown compiler"
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or cores
    threads = min(threads, cores)
    with tempfile.TemporaryDirectory(prefix="ftc_native_") as tmp:
        path_in, path_out = Path(tmp) / "in.bin", Path(tmp) / "out.bin"
        np.concatenate([a.ravel(order="F") for a in initial]).tofile(path_in)
        log(f"bench: native Fortran CPU baseline, {threads} threads")
        reps, median = run_native(path_in, path_out, nx, ny, nz, steps, 7, threads)
        raw = np.fromfile(path_out, dtype=np.float64)
        size = (nx + 2) * (ny + 2) * (nz + 2)
        parity = all(np.array_equal(raw[f * size:(f + 1) * size], g.ravel(order="F")) for f, g in enumerate(first))
        log("bench: native Fortran CPU baseline, 4 threads")
        reps4, median4 = run_native(path_in, None, nx, ny, nz, 1, 7, min(4, cores))
    cpu = {"value": round(median, 3), "unit": "ms/step", "cores": threads, "kind": "port",
           "build": "native Fortran of the same step (oracle/native/les_native.f90): " + compiler,
           "sample": f"full {nx}x{ny}x{nz} grid: 1 warm-up step, then 7 repeats of {steps} steps, median of the "
                     f"repeats; OMP_PROC_BIND=close OMP_PLACES=cores",
           "repeats_ms_per_step": reps, "cpu_model": cpu_model(),
           "threads_4": {"value": round(median4, 3), "repeats_ms_per_step": reps4}}
    return cpu, parity


def main() -> int:
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=100)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--nx", type=int, default=512)
    p.add_argument("--ny", type=int, default=512)
    p.add_argument("--nz", type=int, default=128)
    p.add_argument("--cpu-steps", type=int, default=2,
                   help="steps per timed repeat of the native CPU baseline (0: skip)")
    p.add_argument("--copy-probe", action="store_true",
                   help="only time a plain 2 GiB device copy (HBM ceiling and byte-counter calibration), "
                        "the other HBM rates and the FP64 matrix-core ceiling")
    p.add_argument("--no-extras", action="store_true",
                   help="skip the Poisson (FP64 vs mixed) and chain-contraction sections")
    p.add_argument("--chain-only", action="store_true",
                   help="only time D = A*B*C at N = 4096 (FP64 MFMA and split-bf16) and print its section")
    p.add_argument("--poisson-only", action="store_true",
                   help="only time the FP64 and mixed pressure solves at 512x512x128 and print their section")
    p.add_argument("--ensemble-only", action="store_true",
                   help="only time the 64-member ensemble (per member, batched FP64, batched mixed) and print "
                        "its section")
    p.add_argument("--projected-step-only", action="store_true",
                   help="only time the whole projected step (LES step, divergence, FP64 and mixed solve, "
                        "projection) and print its section")
    args = p.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the LES domain would shard by "
              "x-y slabs with halo exchange; see the design notes)", file=sys.stderr)
        return 2

    ftc = graft.package()
    scope = graft._module("scope")
    fields = graft._module("fields")
    lib = ftc.load()
    if lib.ftc_device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 1

    if args.chain_only:
        section, results, inputs = chain_section(ftc, scope, lib)
        A, B, C = inputs
        ref = (A @ B) @ C
        for name, D in results.items():
            section[name]["max_rel_error_vs_fp64_cpu"] = float(np.abs(D - ref).max() / np.abs(ref).max())
        print(json.dumps({"chain_contraction": section}))
        return 0
    if args.poisson_only:
        print(json.dumps({"poisson": poisson_section(ftc, scope, lib, fields, reps=5)[0]}))
        return 0
    if args.ensemble_only:
        print(json.dumps({"ensemble": ensemble_section(lib, fields)}))
        return 0
    if args.projected_step_only:
        print(json.dumps({"projected_step": projected_step_section(ftc, scope, lib, fields, steps=3)}))
        return 0
    if args.copy_probe:
        n = 1 << 28
        ms = ctypes.c_float()
        scope.check(lib, lib.ftc_copy_probe(n, 20, ctypes.byref(ms)))
        out = {"copy_probe": {"bytes_per_launch": 16 * n, "ms_per_launch": round(ms.value, 4),
                              "GB_per_s": round(16 * n / (ms.value * 1e-3) / 1e9, 1)}}
        for kind, name, factor in ((1, "copy_nontemporal", 16), (2, "read", 8), (3, "write", 8)):
            scope.check(lib, lib.ftc_bandwidth_probe(kind, n, 10, ctypes.byref(ms)))
            out[name] = {"bytes_per_launch": factor * n, "ms_per_launch": round(ms.value, 4),
                         "GB_per_s": round(factor * n / (ms.value * 1e-3) / 1e9, 1)}
        # The FP64 matrix-core ceiling: four-wave workgroups, 1, 2 and 4 per CU (waves per SIMD), eight
        # independent 16x16x4 FP64 MFMAs per step.
        flops = ctypes.c_double()
        for wps in (1, 2, 4):
            scope.check(lib, lib.ftc_mfma_probe(wps, 4000, 3, ctypes.byref(ms), ctypes.byref(flops)))
            out[f"mfma_f64_{wps}_waves_per_simd"] = {
                "ms_per_launch": round(ms.value, 4),
                "TFLOP_per_s": round(flops.value / (ms.value * 1e-3) / 1e12, 1)}
        print(json.dumps(out), flush=True)
        return 0

    nx, ny, nz = args.nx, args.ny, args.nz
    dx = dy = dz = 10.0
    dt, cs = 0.2, 0.17
    log(f"bench: generating fields {nx}x{ny}x{nz}")
    U, V, W, S = fields.les_fields(nx, ny, nz, dx, dy, dz)
    initial = [U.copy(order="F"), V.copy(order="F"), W.copy(order="F"), S.copy(order="F")] if args.cpu_steps else None
    b = [np.zeros_like(U, order="F") for _ in range(4)]
    a = [U, V, W, S]

    dom = scope.Domain(lib)
    ha = [dom.bind(x) for x in a]
    hb = [dom.bind(x) for x in b]
    scalars = [ftc.dp(dx), ftc.dp(dy), ftc.dp(dz), ftc.dp(dt), ftc.dp(cs), ftc.ip(nx), ftc.ip(ny), ftc.ip(nz),
               ftc.ip(1)]

    def step(n: int) -> None:
        src, dst = (ha, hb) if n % 2 == 0 else (hb, ha)
        scope.check(lib, lib.cpp_fort_shared_les_step(dom.ctx, 1, *dst, *src, *scalars))
        dom.flush()

    log("bench: first step (upload + direct launch)")
    step(0)
    dom.observe(*hb)
    first = [x.copy(order="F") for x in b] if args.cpu_steps else None
    for n in range(1, args.warmup):
        step(n)
    dom.wait()

    stream = ctypes.c_void_p(dom.stream())
    ev0, ev1 = ctypes.c_void_p(), ctypes.c_void_p()
    lib.ftc_event_create(ctypes.byref(ev0))
    lib.ftc_event_create(ctypes.byref(ev1))
    log(f"bench: timing {args.steps} steps")
    dom.wait()
    t0 = time.perf_counter()
    lib.ftc_event_record(ev0, stream)
    for n in range(args.warmup, args.warmup + args.steps):
        step(n)
    lib.ftc_event_record(ev1, stream)
    dom.wait()
    wall = time.perf_counter() - t0
    ms = ctypes.c_float()
    lib.ftc_event_elapsed_ms(ev0, ev1, ctypes.byref(ms))
    gpu_ms_per_step = ms.value / args.steps
    wall_ms_per_step = wall * 1e3 / args.steps
    gstats = dom.graph_stats()
    dom.close()

    P = (nx + 2) * (ny + 2) * (nz + 2)
    I = nx * ny * nz
    bytes_alg = 8 * (4 * P + 4 * I)
    achieved = bytes_alg / (gpu_ms_per_step * 1e-3) / 1e9
    # The step's kernel: the 64 x 8 buffer-addressed kernel, its equal-spacing variant when dx = dy = dz (as here).
    traffic = traffic_from_profiles(f"ftc::les_step_kernel<512, 32, {'true' if dx == dy == dz else 'false'}, 4, 3, 2>")

    extras = {}
    if not args.no_extras:
        log("bench: Poisson 512x512x128, FP64 vs mixed")
        extras["poisson"], poisson_sols, poisson_rhs = poisson_section(ftc, scope, lib, fields)
        log("bench: chain contraction 4096")
        log("projected step (LES step + divergence + solve + projection)")
        extras["projected_step"] = projected_step_section(ftc, scope, lib, fields)
        extras["chain_contraction"], chain_results, chain_inputs = chain_section(ftc, scope, lib)
        log("bench: 64-member ensemble")
        extras["ensemble"] = ensemble_section(lib, fields)

    cpu = None
    parity = None
    if args.cpu_steps:
        from oracle import binding as oracle
        from oracle import poisson as poisson_oracle

        if extras:
            log("bench: oracle checks of the Poisson and chain sections")
            want = poisson_oracle.solve(poisson_rhs, (10.0, 10.0, 10.0))
            scale = np.abs(want).max()
            for name, sol in poisson_sols.items():
                extras["poisson"][name]["max_rel_error_vs_oracle"] = float(np.abs(sol - want).max() / scale)
            A, B, C = chain_inputs
            ref = (A @ B) @ C
            for name, D in chain_results.items():
                extras["chain_contraction"][name]["max_rel_error_vs_fp64_cpu"] = float(
                    np.abs(D - ref).max() / np.abs(ref).max())

        cpu, parity = native_baseline(initial, first, nx, ny, nz, args.cpu_steps)

    line = {
        "metric": METRIC,
        "value": round(gpu_ms_per_step, 4),
        "unit": "ms/step",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(gpu_ms_per_step, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"resident LES step: advection + Smagorinsky diffusion (Cs={cs}) + halo fill, "
                               f"FP64, {nx}x{ny}x{nz} cells, one hipGraph replay per step",
                   "nx": nx, "ny": ny, "nz": nz, "dx": dx, "dt": dt},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic},
        "cpu_baseline": cpu,
        "wall_ms_per_step": round(wall_ms_per_step, 4),
        "parity_first_step_bit_exact": parity,
        "graph": {"instantiations": gstats.instantiations, "replays": gstats.replays,
                  "parameter_updates": gstats.parameter_updates},
    }
    line.update(extras)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
