#!/usr/bin/env python3
"""Benchmark of the block-structured stencil path on one MI355X.

One run measures every workload and prints one JSON line.  `value` is the
headline: FP32 GCell-updates/s of the diffusion workload -- 8x8x8 blocks of
32^3 cells (256^3), periodic, ghost width 1, seeded sum of 8 Fourier modes,
dt = 0.1 h^2, 100 steps, one "step" = lab + 7-point Laplacian + forward Euler
+ FP64 L2 norm of the new field over the whole grid (one diffusion_run,
two steps per launch on this periodic grid), in FP32 and in FP64 from the
same initial field.  `workloads` holds one record each for the other
workloads at their full sizes: WENO5 face states (512^3 x 5 components),
the stretched-z Jacobi solve (512^3 of 64^3 blocks, FP32 sweeps with FP64
refinement) and the batch of 96 uneven grids.  Every record carries its
FP32-vs-FP64 error (or its result check), the roofline of its dominant
kernel (HIP events on the launch stream; the bytes model is stated in the
record) and a CPU baseline (the C++ oracle on the host cores and on one
core, CPU model stated).

  python bench.py [--steps K] [--warmup W]
                  [--workload all|diffusion|weno5|jacobi|batch|copy]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GCell-updates/s + HBM GB/s % peak, FP32 vs FP64 path at stated rel-L2 error"
HBM_PEAK_GBS = 8000.0


class HipEvents:
    """hipEvent timing on a given stream (the stream the kernels run on)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipDeviceSynchronize.argtypes = []

    def new(self):
        e = ctypes.c_void_p()
        assert self.hip.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.hip.hipEventRecord(e, stream.handle) == 0

    def elapsed_ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        ms = ctypes.c_float()
        assert self.hip.hipEventElapsedTime(ctypes.byref(ms), a, b) == 0
        return float(ms.value)

    def device_sync(self):
        assert self.hip.hipDeviceSynchronize() == 0


def roofline(bytes_per_launch, avg_s, traffic=None, bytes_model=None):
    ach = bytes_per_launch / avg_s / 1e9
    r = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic}
    if bytes_model:
        r["bytes_model"] = bytes_model
    return r


# ------------------------------------------------------------------ CPU baseline


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """Host threads for the all-core baseline: the job's CPU share
    (OMP_NUM_THREADS when the environment sets it), else every core this
    process may run on."""
    n = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def cpu_timed(fn, budget_s, min_samples=10, max_samples=10000):
    """One warm-up call, then samples until at least min_samples and the
    budget are both reached (the protocol of ISPC/BaseBenchmark.cpp:65-165:
    warm-up, repeated samples, mean / min / sdev); a sample that alone
    exceeds the budget ends the run after three samples."""
    fn()
    times = []
    t_start = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
        el = time.perf_counter() - t_start
        if len(times) >= max_samples:
            break
        if len(times) >= min_samples and el >= budget_s:
            break
        if len(times) >= 3 and el >= 3 * budget_s:
            break
    a = np.array(times)
    return {"samples": int(a.size), "mean_s": float(a.mean()), "min_s": float(a.min()), "sdev_s": float(a.std())}


def cpu_baseline(work_all, cells_all, sample_all, work_one, cells_one, sample_one, budget_s):
    """CPU baseline record: work_all(nthreads) on the job's host cores and
    work_one() on one core, each timed by cpu_timed; `cells_*` are the
    cell-updates of one call.  value = cells / mean sample time."""
    nt = cpu_threads()
    a = cpu_timed(lambda: work_all(nt), budget_s)
    o = cpu_timed(work_one, budget_s)
    return {"value": round(cells_all / a["mean_s"] / 1e9, 5), "unit": "GCell-updates/s", "cores": nt,
            "kind": "port", "sample": sample_all, "cpu_model": cpu_model(), "os_cpu_count": os.cpu_count(),
            "timing": a, "best_value": round(cells_all / a["min_s"] / 1e9, 5),
            "one_core": {"value": round(cells_one / o["mean_s"] / 1e9, 5), "unit": "GCell-updates/s", "cores": 1,
                         "sample": sample_one, "timing": o}}


def measured_traffic(name, key):
    """HBM bytes per launch of one kernel from a committed rocprofv3 --pmc
    summary: the line `hbm_bytes_per_launch[key] = N` of profiles/<name>
    (FETCH_SIZE and WRITE_SIZE scaled by their copy-kernel calibration)."""
    import re
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    m = re.search(r"^hbm_bytes_per_launch\[%s\] = (\d+)" % re.escape(key), open(p).read(), re.M)
    return int(m.group(1)) if m else None


# ------------------------------------------------------------------ diffusion


def run_launches(steps):
    """Launches of a diffusion_run of `steps` steps on a periodic grid of
    32^3 blocks: two-step launches up to a multiple of four, then one-step."""
    fused = 2 * (steps // 4)
    return fused + (steps - 2 * fused)


def diffusion_gpu(cnv, ev, dtype, u0, steps, warmup, autotune=True, tile=None, tile2=None):
    nb, n, N = (8, 8, 8), (32, 32, 32), 256
    h = 1.0 / N
    dt = 0.1 * h * h
    a = cnv.Cartesian(nb, n, dtype=dtype)
    b = cnv.Cartesian(nb, n, dtype=dtype)
    a.upload_global(0, u0.astype(dtype))
    tune = None
    if tile:
        cnv.set_tune_stencil7(dtype, 32, tile[0], tile[1])
        tune = {"wg": tile[0], "zchunk": tile[1]}
    elif autotune:
        wg, zc, table = cnv.autotune_stencil7(a, b, iters=10)
        tune = {"wg": wg, "zchunk": zc}
        print("autotune %s: best wg=%d zchunk=%d; %s" % (np.dtype(dtype).name, wg, zc,
              " ".join("%d/%d:%.1fus" % t for t in table)), file=sys.stderr, flush=True)
        a.upload_global(0, u0.astype(dtype))
    # the two-steps-per-launch kernel that diffusion_run uses on this periodic grid
    tune2 = None
    if tile2:
        cnv.set_tune_diffusion_run(dtype, 32, tile2[0], tile2[1])
        tune2 = {"wg": tile2[0], "zchunk": tile2[1]}
    elif autotune:
        wg, zc, table = cnv.autotune_diffusion_run(a, b, iters=10)
        tune2 = {"wg": wg, "zchunk": zc}
        print("autotune two-step %s: best wg=%d zchunk=%d; %s" % (np.dtype(dtype).name, wg, zc,
              " ".join("%d/%d:%.1fus" % t for t in table)), file=sys.stderr, flush=True)
        a.upload_global(0, u0.astype(dtype))
    else:
        wg, zc = cnv.tune_diffusion_run(dtype, 32)
        tune2 = {"wg": wg, "zchunk": zc}
    st = cnv.Stream()
    norms = cnv.DeviceBuffer(8 * max(1, steps + warmup))
    a.prepare()
    b.prepare()
    src, dst = a, b
    if warmup:
        res = cnv.diffusion_run(src, dst, warmup, dt, norms=norms, stream=st)
        if res is dst:
            src, dst = dst, src
    st.synchronize()
    # record exactly `steps` steps (one multi-step run, each step's norm
    # reduced inside the next step's launch) into one graph
    g = cnv.Graph(st).capture()
    s2 = cnv.diffusion_run(src, dst, steps, dt, norms=norms.at(8 * warmup), stream=st)
    g.end()
    # the graph's first launch carries one-time set-up: one untimed replay
    # (the same steps in both precisions), then one timed replay of the
    # `steps` steps: the wall clock between two device synchronisations, the
    # launches' own time from HIP events on the stream
    g.launch()
    st.synchronize()
    e0, e1 = ev.new(), ev.new()
    ev.device_sync()
    t0 = time.perf_counter()
    ev.record(e0, st)
    g.launch()
    ev.record(e1, st)
    ev.device_sync()
    wall = time.perf_counter() - t0
    kern_ms = ev.elapsed_ms(e0, e1)
    final = s2.download(0)
    hist = norms.download(np.float64, steps + warmup)
    cells = 256 ** 3
    return {
        "wall_s": wall, "event_ms": kern_ms, "final": final, "norms": hist,
        "gcups": cells * steps / wall / 1e9,
        # every launch (two-step or one-step) reads the field once and writes it once
        "bytes_per_launch": cells * 2 * np.dtype(dtype).itemsize,
        "launches": run_launches(steps), "avg_kernel_s": kern_ms / 1e3 / run_launches(steps),
        "tile": tune, "tile2": tune2,
    }


def cpu_baseline_diffusion(u0, budget_s):
    import oracle   # the CPU baseline leg only
    from cubismnova_amd.layout import global_to_blocks
    nb, n, N = (8, 8, 8), (32, 32, 32), 256
    f = oracle.Field(global_to_blocks(u0, nb, n), nb, n)
    h = [1.0 / N] * 3
    dt = 0.1 / N / N
    return cpu_baseline(lambda nt: oracle.diffusion_step(f, h, dt, nthreads=nt), N ** 3,
                        "one FP64 diffusion step of the full 256^3 grid per sample (C++ oracle: per-block lab + "
                        "7-point Laplacian in the Laplacian.cpp order + two-rounding Euler + FP64 norm, OpenMP "
                        "over blocks with thread-local labs)",
                        lambda: oracle.diffusion_step(f, h, dt, nthreads=1), N ** 3,
                        "the same step on one thread (the per-block Laplacian.cpp loop)", budget_s)


def bench_diffusion(args):
    import cubismnova_amd as cnv
    from cubismnova_amd import workloads
    ev = HipEvents()
    u0 = workloads.diffusion_large(256)
    r32 = diffusion_gpu(cnv, ev, np.float32, u0, args.steps, args.warmup, not args.no_autotune, args.tile_f32,
                        args.tile2_f32)
    r64 = diffusion_gpu(cnv, ev, np.float64, u0, args.steps, args.warmup, not args.no_autotune, args.tile_f64,
                        args.tile2_f64)
    f32, f64 = r32["final"].astype(np.float64), r64["final"]
    rel = float(np.linalg.norm(f32 - f64) / np.linalg.norm(f64))
    nrel = float(np.max(np.abs(r32["norms"] - r64["norms"]) / r64["norms"]))

    def rl(r, isz, key):
        # one launch = two steps (one read of u^n, one write of u^{n+2}); the
        # per-step model of 2*sizeof(T) per cell-update counts that pass once
        # per step, so its rate is the launch rate times the steps per launch
        out = roofline(r["bytes_per_launch"], r["avg_kernel_s"], measured_traffic("stencil7_diffusion_256.md", key),
                       "%d B per cell per launch (u read once, u written once); a launch advances %.2f steps on "
                       "average over the run (two-step launches, single steps for the remainder)"
                       % (2 * isz, args.steps / r["launches"]))
        per_step = 2 * isz * 256 ** 3 * args.steps / (r["avg_kernel_s"] * r["launches"]) / 1e9
        out["per_step_model"] = {"bytes_per_cell_update": 2 * isz, "achieved": round(per_step, 1),
                                 "frac": round(per_step / HBM_PEAK_GBS, 4),
                                 "note": "counts one read and one write per step; above the peak only because "
                                         "two steps share one pass over HBM"}
        return out

    out = {
        "metric": METRIC, "value": round(r32["gcups"], 3), "unit": "GCell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(r32["wall_s"] * 1e3 / args.steps, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded sum of 8 Fourier modes, |k| in [1,16], amplitude 1/|k|)",
        "config": {"workload": "diffusion: 8x8x8 blocks of 32^3 cells (256^3), periodic, ghost 1, fused "
                               "lab + 7-point Laplacian + forward Euler (dt=0.1h^2) + FP64 L2 norm per step, "
                               "two steps per launch",
                   "blocks": [8, 8, 8], "block_cells": [32, 32, 32], "steps_per_graph": args.steps,
                   "launches_per_graph": r32["launches"], "tile_f32": r32["tile"], "tile_f64": r64["tile"],
                   "tile2_f32": r32["tile2"], "tile2_f64": r64["tile2"]},
        "roofline": rl(r32, 4, "f32"),
        "kernel_ms_per_step": round(r32["event_ms"] / args.steps, 5),
        "fp64": {"value": round(r64["gcups"], 3), "unit": "GCell-updates/s",
                 "ms_per_step": round(r64["wall_s"] * 1e3 / args.steps, 5),
                 "kernel_ms_per_step": round(r64["event_ms"] / args.steps, 5),
                 "roofline": rl(r64, 8, "f64")},
        "fp32_over_fp64": round(r32["gcups"] / r64["gcups"], 3),
        "fp32_rel_l2_vs_fp64": rel,
        "fp32_norm_history_max_rel_dev": nrel,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_diffusion(u0, args.cpu_seconds)
    return out


# ------------------------------------------------------------------ WENO5


def weno_gpu(cnv, ev, dtype, steps, warmup, autotune, comps, tile=None):
    N, nb, n = 512, (16, 16, 16), (32, 32, 32)
    g = cnv.Cartesian(nb, n, dtype=dtype, ncomp=5)
    for c in range(5):
        g.upload_global(c, comps[c].astype(dtype))
    for bc in (cnv.Absorbing(1, 0), cnv.Absorbing(1, 1), cnv.Absorbing(2, 0), cnv.Absorbing(2, 1)):
        g.addFaceBC(bc)              # x periodic, y/z zero-gradient
    out = cnv.Cartesian(nb, n, dtype=dtype, ncomp=30)
    tune = None
    if tile:
        cnv.set_tune_weno5(dtype, tile[0], tile[1])
        tune = {"wg": tile[0], "zchunk": tile[1]}
    elif autotune:
        wg, zc, table = cnv.autotune_weno5(g, out, iters=3)
        tune = {"wg": wg, "zchunk": zc}
        print("autotune weno5 %s: best wg=%d zchunk=%d; %s" % (np.dtype(dtype).name, wg, zc,
              " ".join("%d/%d:%.0fus" % t for t in table)), file=sys.stderr, flush=True)
    st = cnv.Stream()
    for _ in range(warmup):
        cnv.weno5(g, out, stream=st)
    st.synchronize()
    gr = cnv.Graph(st).capture()
    for _ in range(steps):
        cnv.weno5(g, out, stream=st)
    gr.end()
    gr.launch()   # the graph's first launch carries one-time set-up: untimed
    st.synchronize()
    e0, e1 = ev.new(), ev.new()
    ev.device_sync()
    t0 = time.perf_counter()
    ev.record(e0, st)
    gr.launch()
    ev.record(e1, st)
    st.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = ev.elapsed_ms(e0, e1)
    cells = N ** 3
    return {"grid": g, "out": out, "wall_s": wall, "gcups": cells * steps / wall / 1e9,
            "bytes_per_launch": cells * 35 * np.dtype(dtype).itemsize, "avg_kernel_s": kern_ms / 1e3 / steps,
            "tile": tune}


def cpu_baseline_weno(grid64, budget_s):
    import oracle   # the CPU baseline leg only
    nb, n = grid64.nblocks, grid64.block_cells
    cells_blk = int(np.prod(n))
    slab = grid64.download(0)
    bcl = {}
    for b in range(grid64.nblocks_total):
        by, bz = (b // nb[0]) % nb[1], b // (nb[0] * nb[1])
        lst = []
        if by == 0:
            lst.append((1, 1, 0, 0.0))
        if by == nb[1] - 1:
            lst.append((1, 1, 1, 0.0))
        if bz == 0:
            lst.append((1, 2, 0, 0.0))
        if bz == nb[2] - 1:
            lst.append((1, 2, 1, 0.0))
        bcl[b] = lst
    f = oracle.Field(slab, nb, n, grid64.block_stride, bcl)
    nt = cpu_threads()
    # spread the sampled blocks over the grid (corner, face and interior blocks)
    sel_all = [(j * 997) % grid64.nblocks_total for j in range(4 * nt)]
    sel_one = sel_all[:4]
    # one cell-update = all 5 components; a sample runs one component
    return cpu_baseline(lambda t: oracle.weno5(f, prec=1, nthreads=t, blocks=sel_all), len(sel_all) * cells_blk / 5,
                        "x/y/z WENO5 states of %d blocks of 32^3 of one component per sample (FP64 C++ oracle, "
                        "per-block labs, OpenMP over blocks), counted as 1/5 of a cell-update per cell"
                        % len(sel_all),
                        lambda: oracle.weno5(f, prec=1, nthreads=1, blocks=sel_one), len(sel_one) * cells_blk / 5,
                        "the same on 4 blocks, one thread", budget_s)


def bench_weno5(args, steps, warmup):
    import cubismnova_amd as cnv
    from cubismnova_amd import workloads
    ev = HipEvents()
    comps = [workloads.weno_component(512, c) for c in range(5)]
    r32 = weno_gpu(cnv, ev, np.float32, steps, warmup, not args.no_autotune, comps, args.tile_f32)
    r64 = weno_gpu(cnv, ev, np.float64, steps, warmup, not args.no_autotune, comps, args.tile_f64)
    del comps
    worst = 0.0
    for k in range(30):
        d2, b2, _ = cnv.grid_compare(r32["out"], r64["out"], k, k)
        worst = max(worst, float(np.sqrt(d2 / b2)))
    model = "%d B per cell per launch: 5 components read once, 30 face states (2 per direction and component) " \
            "written once"
    out = {
        "metric": METRIC, "value": round(r32["gcups"], 3), "unit": "GCell-updates/s",
        "steps": steps, "warmup": warmup, "ms_per_step": round(r32["wall_s"] * 1e3 / steps, 4),
        "kernel_ms_per_step": round(r32["avg_kernel_s"] * 1e3, 4),
        "dtype": "f32", "data": "synthetic (smooth compressible state + tanh pressure front, ratio 10, seeded)",
        "config": {"workload": "WENO5 face states in x, y, z: 16x16x16 blocks of 32^3 (512^3), 5 components SoA, "
                               "ghost 3, x periodic, y/z absorbing", "blocks": [16, 16, 16],
                   "block_cells": [32, 32, 32], "components": 5, "tile_f32": r32["tile"], "tile_f64": r64["tile"]},
        "roofline": roofline(r32["bytes_per_launch"], r32["avg_kernel_s"], measured_traffic("weno5_512.md", "f32"),
                             model % 140),
        "fp64": {"value": round(r64["gcups"], 3), "unit": "GCell-updates/s",
                 "ms_per_step": round(r64["wall_s"] * 1e3 / steps, 4),
                 "kernel_ms_per_step": round(r64["avg_kernel_s"] * 1e3, 4),
                 "roofline": roofline(r64["bytes_per_launch"], r64["avg_kernel_s"],
                                      measured_traffic("weno5_512.md", "f64"), model % 280)},
        "fp32_over_fp64": round(r32["gcups"] / r64["gcups"], 3),
        "fp32_rel_l2_vs_fp64_worst_state": worst,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_weno(r64["grid"], args.cpu_seconds)
    return out


# ------------------------------------------------------------------ Jacobi


def jacobi_grids(cnv, N, nb, n, dtype, zf, ncomp=1):
    g = cnv.Cartesian(nb, n, dtype=dtype, ncomp=ncomp)
    for d in range(3):
        for s in (0, 1):
            g.addFaceBC(cnv.Dirichlet(d, s, 0.0))
    cnv.set_zfaces(g, zf)
    return g


def bench_jacobi(args, steps, warmup):
    import cubismnova_amd as cnv
    from cubismnova_amd import stretch, workloads
    ev = HipEvents()
    N, nb, n = 512, (8, 8, 8), (64, 64, 64)
    zf = stretch.tanh_faces(N, 2.0)
    czm, czp, dz, zc = stretch.metric(zf)
    src = workloads.poisson_source(N, zc)
    u = jacobi_grids(cnv, N, nb, n, np.float64, zf)
    f = jacobi_grids(cnv, N, nb, n, np.float64, zf)
    u.upload_global(0, np.zeros((N, N, N)))
    f.upload_global(0, src)
    work = jacobi_grids(cnv, N, nb, n, np.float32, zf, ncomp=3)
    S = 50
    # sweep tiles: timed on the workload's own grids per dtype (or the built-in table)
    tiles = {}
    for dt in (np.float32, np.float64):
        key = np.dtype(dt).name
        pin = args.tile_f32 if dt == np.float32 else args.tile_f64
        if pin:
            cnv.set_tune_jacobi(dt, 64, pin[0], pin[1], pin[2] if len(pin) > 2 else 0)
        elif not args.no_autotune:
            ga, gb, gf = (jacobi_grids(cnv, N, nb, n, dt, zf) for _ in range(3))
            ga.upload_global(0, np.zeros((N, N, N), dtype=dt))
            gf.upload_global(0, src.astype(dt))
            tw, tz, tp, table = cnv.autotune_jacobi(ga, gf, gb, iters=5)
            print("autotune Jacobi sweep %s: best wg=%d zchunk=%d prefetch=%d; %s" % (
                key, tw, tz, tp, " ".join("%d/%d/%d:%.1fus" % t for t in table)), file=sys.stderr, flush=True)
            del ga, gb, gf
        tw, tz, tp = cnv.tune_jacobi(dt, 64)
        tiles[key] = {"wg": tw, "zchunk": tz, "f_prefetch": tp}
    # warm: one refined cycle, then `steps` cycles timed (each = 50 FP32 sweeps
    # + FP64 correction + FP64 defect, one graph replay per cycle)
    cnv.jacobi_solve(u, f, work, refine=True, sweeps_per_cycle=S, max_cycles=max(1, warmup), tol=0.0)
    ev.device_sync()
    t0 = time.perf_counter()
    cyc, sw, rel, hist = cnv.jacobi_solve(u, f, work, refine=True, sweeps_per_cycle=S, max_cycles=steps,
                                          tol=0.0)
    ev.device_sync()
    wall = time.perf_counter() - t0
    cells = N ** 3
    # the FP32 sweeps alone as the solver runs them (work comps 0 <-> 1, rhs
    # comp 2): S / 2 two-sweep launches, then S single sweeps; HIP events on
    # their stream
    st = cnv.Stream()

    def graph_time(launch, count):
        gr = cnv.Graph(st).capture()
        for k in range(count):
            launch(k)
        gr.end()
        gr.launch()
        st.synchronize()
        e0, e1 = ev.new(), ev.new()
        ev.record(e0, st)
        gr.launch()
        ev.record(e1, st)
        return ev.elapsed_ms(e0, e1) / 1e3

    pair_s = graph_time(lambda k: cnv.jacobi_sweep2(work, work, work, 2.0 / 3.0, cu=k % 2, cf=2, co=1 - k % 2,
                                                   stream=st), S // 2) / (S // 2)
    single_s = graph_time(lambda k: cnv.jacobi_sweep(work, work, work, 2.0 / 3.0, cu=k % 2, cf=2, co=1 - k % 2,
                                                     stream=st), S) / S
    sweep_s = pair_s / 2
    # FP64 sweeps (all-FP64 Jacobi) on the same grid
    u2 = jacobi_grids(cnv, N, nb, n, np.float64, zf)
    w2 = jacobi_grids(cnv, N, nb, n, np.float64, zf)
    cnv.call("cnv_fill", u2.data_ptr(0), 1, 0.0, u2.component_bytes // 8, None)
    u2.prepare()
    w2.prepare()
    pair64_s = graph_time(lambda k: cnv.jacobi_sweep2(*((u2, f, w2) if k % 2 == 0 else (w2, f, u2)), 2.0 / 3.0,
                                                      stream=st), S // 2) / (S // 2)
    single64_s = graph_time(lambda k: cnv.jacobi_sweep(*((u2, f, w2) if k % 2 == 0 else (w2, f, u2)), 2.0 / 3.0,
                                                       stream=st), S) / S
    sweep64_s = pair64_s / 2
    tiles["float32"]["two_sweep"] = list(cnv.get_tune_jacobi2(np.float32, 64))
    tiles["float64"]["two_sweep"] = list(cnv.get_tune_jacobi2(np.float64, 64))
    # convergence to 1e-12 at 32^3: refined vs all-FP64 Jacobi vs direct solve
    Ns, nbs, ns = 32, (2, 2, 2), (16, 16, 16)
    zfs = stretch.tanh_faces(Ns, 2.0)
    _, _, _, zcs = stretch.metric(zfs)
    srcs = workloads.poisson_source(Ns, zcs)
    us, fs_ = jacobi_grids(cnv, Ns, nbs, ns, np.float64, zfs), jacobi_grids(cnv, Ns, nbs, ns, np.float64, zfs)
    us.upload_global(0, np.zeros((Ns,) * 3))
    fs_.upload_global(0, srcs)
    c1, s1, r1, _ = cnv.jacobi_solve(us, fs_, jacobi_grids(cnv, Ns, nbs, ns, np.float32, zfs, 3), refine=True,
                                     sweeps_per_cycle=S, max_cycles=20000, tol=1e-12)
    u64 = jacobi_grids(cnv, Ns, nbs, ns, np.float64, zfs)
    u64.upload_global(0, np.zeros((Ns,) * 3))
    c2, s2, r2, _ = cnv.jacobi_solve(u64, fs_, jacobi_grids(cnv, Ns, nbs, ns, np.float64, zfs, 2), refine=False,
                                     sweeps_per_cycle=S, max_cycles=20000, tol=1e-12)
    d2, b2, mx = cnv.grid_compare(us, u64)
    out = {
        "metric": METRIC, "value": round(cells / sweep_s / 1e9, 3), "unit": "GCell-updates/s",
        "steps": steps, "warmup": warmup, "ms_per_step": round(wall * 1e3 / steps, 4),
        "dtype": "f32", "data": "synthetic (3 seeded Gaussian sources, sigma 0.05, mean removed)",
        "config": {"workload": "iteratively refined weighted Jacobi (omega 2/3): 8x8x8 blocks of 64^3 (512^3), "
                               "tanh-stretched z (factor 2), Dirichlet 0; a step = one cycle of 50 FP32 sweeps "
                               "+ FP64 correction + FP64 defect", "sweeps_per_cycle": S,
                   "value_is": "FP32 sweep GCell-updates/s (cells per sweep / sweep time), sweeps run as "
                               "the solver runs them: two per launch",
                   "tile_f32": tiles["float32"], "tile_f64": tiles["float64"]},
        "kernel_ms_per_sweep": round(sweep_s * 1e3, 4),
        "kernel_ms_per_launch": round(pair_s * 1e3, 4),
        "sweeps_per_s": round(sw / wall, 1), "cycle_gcups": round(cells * sw / wall / 1e9, 3),
        "roofline": roofline(cells * 12, pair_s, measured_traffic("jacobi_sweep_512.md", "f32x2"),
                             "12 B per cell per FP32 launch of two sweeps: u and f read once, u'' written once "
                             "(the z metric is two 1-D arrays)"),
        "single_sweep": {"value": round(cells / single_s / 1e9, 3), "unit": "GCell-updates/s",
                         "kernel_ms_per_sweep": round(single_s * 1e3, 4),
                         "roofline": roofline(cells * 12, single_s, measured_traffic("jacobi_sweep_512.md", "f32"),
                                              "12 B per cell per FP32 sweep")},
        "fp64": {"value": round(cells / sweep64_s / 1e9, 3), "unit": "GCell-updates/s",
                 "kernel_ms_per_sweep": round(sweep64_s * 1e3, 4),
                 "kernel_ms_per_launch": round(pair64_s * 1e3, 4),
                 "roofline": roofline(cells * 24, pair64_s, measured_traffic("jacobi_sweep_512.md", "f64x2"),
                                      "24 B per cell per FP64 launch of two sweeps"),
                 "single_sweep": {"value": round(cells / single64_s / 1e9, 3),
                                  "kernel_ms_per_sweep": round(single64_s * 1e3, 4),
                                  "roofline": roofline(cells * 24, single64_s,
                                                       measured_traffic("jacobi_sweep_512.md", "f64"),
                                                       "24 B per cell per FP64 sweep")}},
        "fp32_over_fp64": round(sweep64_s / sweep_s, 3),
        "convergence_32cube": {"refined_sweeps": int(s1), "refined_rel_residual": r1, "fp64_sweeps": int(s2),
                               "fp64_rel_residual": r2,
                               "refined_vs_fp64_rel_l2": float(np.sqrt(d2 / b2)),
                               "refined_vs_fp64_max": mx},
        "residual_after_timed_cycles": rel,
    }
    if not args.no_cpu:
        import oracle   # the CPU baseline leg only

        def dirichlet_lists(nbk):
            lists = {}
            for b in range(nbk[0] * nbk[1] * nbk[2]):
                bi = (b % nbk[0], (b // nbk[0]) % nbk[1], b // (nbk[0] * nbk[1]))
                lst = [(2, d, sd, 0.0) for d in range(3) for sd in (0, 1) if bi[d] == (0 if sd == 0 else nbk[d] - 1)]
                if lst:
                    lists[b] = lst
            return lists

        fo = oracle.Field(u.download(0), nb, n, u.block_stride, dirichlet_lists(nb))
        fsl = f.download(0)
        # one-core sample: a 4x4x4-block (256^3) grid of the same blocks,
        # sources and z metric
        nb1 = (4, 4, 4)
        bs = u.block_stride
        fo1 = oracle.Field(np.zeros(64 * bs), nb1, n, bs, dirichlet_lists(nb1))
        fs1 = np.ascontiguousarray(fsl[:64 * bs])
        out["cpu_baseline"] = cpu_baseline(
            lambda t: oracle.jacobi(fo, fsl, czm, czp, [1.0 / N] * 3, 2.0 / 3.0, 0, nthreads=t), cells,
            "one FP64 weighted-Jacobi sweep of the full 512^3 stretched grid per sample (C++ oracle, per-block "
            "labs, OpenMP over blocks)",
            lambda: oracle.jacobi(fo1, fs1, czm[:256], czp[:256], [1.0 / N] * 3, 2.0 / 3.0, 0, nthreads=1),
            256 ** 3, "one FP64 sweep of a 4x4x4-block (256^3) grid of the same 64^3 blocks, one thread",
            args.cpu_seconds)
    return out


# ------------------------------------------------------------------ batch


def bench_batch(args, steps, warmup):
    import concurrent.futures
    import cubismnova_amd as cnv
    from cubismnova_amd import workloads
    ev = HipEvents()
    specs = workloads.batch_specs(96, k=4)
    # seeded fields of every grid, generated on host threads
    with concurrent.futures.ThreadPoolExecutor(max_workers=cpu_threads()) as ex:
        fields = list(ex.map(lambda sc: [workloads.batch_field(sc[0], c).astype(np.float32)
                                         for c in range(sc[0]["ncomp"])], [(sp,) for sp in specs]))
    grids, items = [], []
    cells_total, bytes_total = 0, 0
    for s, fl in zip(specs, fields):
        nb, e = s["nblocks"], s["edge"]
        g = cnv.Cartesian(nb, e, dtype=np.float32, ncomp=s["ncomp"])
        for c in range(s["ncomp"]):
            g.upload_global(c, fl[c])
        for kind, d, side, v in s["bcs"]:
            g.addFaceBC(cnv.Dirichlet(d, side, v) if kind == "dirichlet" else cnv.Absorbing(d, side))
        nout = 6 * s["ncomp"] if s["op"] == "weno5" else s["ncomp"]
        o = cnv.Cartesian(nb, e, dtype=np.float32, ncomp=nout)
        h = 1.0 / (nb[0] * e)
        items.append((g, o, s["op"], s["ghost"], 0.1 * h * h))
        grids.append((g, o, s))
        cells = nb[0] * nb[1] * nb[2] * e ** 3
        cells_total += cells
        bytes_total += cells * s["ncomp"] * (8 if s["op"] == "diffusion" else 28)
    b = cnv.Batch(items)
    table = [] if args.no_autotune else b.autotune(iters=3)
    if table:
        print("autotune batch (op/edge/wg/depth): " + " ".join("%d/%d/%d/%d:%.0fus" % t for t in table),
              file=sys.stderr, flush=True)
    # the forms as the autotuner timed them (whole runs)
    form_us = {cnv.Batch.FORM_NAMES[int(t[2])]: round(t[4], 1) for t in table if t[0] == -3}
    if args.batch_form != "auto":
        b.set_form({"one": cnv.Batch.SINGLE, "classes": cnv.Batch.CLASSES,
                    "serial": cnv.Batch.CLASSES_SERIAL}[args.batch_form])
    form, nunits = b.form()
    one = form == cnv.Batch.SINGLE
    del fields
    st = cnv.Stream()
    for _ in range(warmup):
        b.run(stream=st)
    st.synchronize()
    gr = cnv.Graph(st).capture()
    for _ in range(steps):
        b.run(stream=st)
    gr.end()
    gr.launch()   # the graph's first launch carries one-time set-up: untimed
    st.synchronize()
    e0, e1 = ev.new(), ev.new()
    ev.device_sync()
    t0 = time.perf_counter()
    ev.record(e0, st)
    gr.launch()
    ev.record(e1, st)
    st.synchronize()
    wall = time.perf_counter() - t0
    kern_s = ev.elapsed_ms(e0, e1) / 1e3 / steps
    info = b.info()
    # after the timed replay: every grid against its single-grid entry point
    # (diffusion bit for bit, WENO5 states within 1e-6 of their largest value)
    worst = 0.0
    for g, o, s in grids:
        ref = cnv.Cartesian(g.nblocks, g.block_cells, dtype=np.float32, ncomp=o.ncomp)
        if s["op"] == "diffusion":
            h = 1.0 / (g.nblocks[0] * g.block_cells[0])
            cnv.diffusion_step(g, ref, 0.1 * h * h)
        else:
            cnv.weno5(g, ref)
        for c in range(o.ncomp):
            x, y = o.download(c).astype(np.float64), ref.download(c).astype(np.float64)
            if s["op"] == "weno5":
                worst = max(worst, float(np.max(np.abs(x - y))) / max(float(np.max(np.abs(y))), 1e-300))
            elif not np.array_equal(x, y):
                worst = 1.0
        del ref
    out = {
        "metric": METRIC, "value": round(cells_total * steps / wall / 1e9, 3), "unit": "GCell-updates/s",
        "steps": steps, "warmup": warmup, "ms_per_step": round(wall * 1e3 / steps, 4),
        "kernel_ms_per_step": round(kern_s * 1e3, 4),
        "dtype": "f32", "data": "synthetic (seeded Fourier-mode fields per grid)",
        "config": {"workload": "96 independent grids (blocks 1x1x1..8x8x4, edges {8,16,32}, 1-3 components, "
                               "ghost 1 or 3, per-face periodic/Dirichlet/absorbing), each a diffusion step or "
                               "WENO5, HIP-graph replayed; the step runs in the form the autotuner timed faster: "
                               "one persistent launch over a device list of work units, or one launch per "
                               "(operator, edge) class of 16^3/32^3 blocks plus a persistent launch for the rest",
                   "form": cnv.Batch.FORM_NAMES[form], "form_times_us": form_us,
                   "units": nunits,
                   "grids": len(specs), "cells": cells_total, "work_items": info["items"],
                   "classes": [dict(zip(("op", "edge", "wg", "zdepth", "items"), c)) for c in b.classes()],
                   "persistent_wg": info["wg"], "persistent_lds_bytes": info["lds_bytes"]},
        "roofline": roofline(bytes_total, kern_s,
                             measured_traffic("batch_96.md", "batch_single_launch" if one else "batch_class_launches"),
                             "sum over grids of cells x components x (8 B for a diffusion step: one read, one "
                             "write; 28 B for WENO5: one read, six states written)"),
        "batch_vs_single_grid_max_rel": worst,
    }
    if not args.no_cpu:
        import oracle   # the CPU baseline leg only
        host = []   # the oracle's inputs, staged before the clock starts
        for g, o, s in grids:
            nb, e = s["nblocks"], s["edge"]
            lists = {}
            for kind, d, side, v in s["bcs"]:
                for bb in range(g.nblocks_total):
                    bi = (bb % nb[0], (bb // nb[0]) % nb[1], bb // (nb[0] * nb[1]))
                    if bi[d] == (0 if side == 0 else nb[d] - 1):
                        lists.setdefault(bb, []).append((2 if kind == "dirichlet" else 1, d, side, v))
            host.append((s, g.nblocks_total * e ** 3,
                         [oracle.Field(g.download(c), nb, e, g.block_stride, lists) for c in range(s["ncomp"])]))

        def one_pass(sel, nt):
            for s, cells, fl in sel:
                nb, e = s["nblocks"], s["edge"]
                for fo in fl:
                    if s["op"] == "diffusion":
                        oracle.diffusion_step(fo, [1.0 / (nb[d] * e) for d in range(3)], 1e-6, nthreads=nt)
                    else:
                        oracle.weno5(fo, prec=1, nthreads=nt)

        sub = host[::8]
        out["cpu_baseline"] = cpu_baseline(
            lambda t: one_pass(host, t), cells_total,
            "one full pass over the 96 grids per sample (%d cells; FP64 C++ oracle, OpenMP over blocks)"
            % cells_total,
            lambda: one_pass(sub, 1), sum(c for _, c, _ in sub),
            "a pass over every 8th grid (%d grids, %d cells), one thread" % (len(sub), sum(c for _, c, _ in sub)),
            args.cpu_seconds)
    return out


# ------------------------------------------------------------------ calibration


def bench_copy(args):
    """Device-to-device copies of a known size: the HBM counters (FETCH_SIZE,
    WRITE_SIZE) read for this kernel calibrate the counters of the others."""
    import cubismnova_amd as cnv
    from cubismnova_amd import _abi
    ev = HipEvents()
    nbytes = 256 << 20
    a, b = cnv.DeviceBuffer(nbytes), cnv.DeviceBuffer(nbytes)
    a.fill(np.float32, 1.0)
    st = cnv.Stream()
    for _ in range(args.warmup):
        _abi.call("cnv_memcpy_d2d", b.ptr, a.ptr, nbytes, st.handle)
    st.synchronize()
    e0, e1 = ev.new(), ev.new()
    ev.record(e0, st)
    for _ in range(args.steps):
        _abi.call("cnv_memcpy_d2d", b.ptr, a.ptr, nbytes, st.handle)
    ev.record(e1, st)
    avg = ev.elapsed_ms(e0, e1) / 1e3 / args.steps
    return {"metric": "device copy GB/s (counter calibration)", "value": round(2 * nbytes / avg / 1e9, 1),
            "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(avg * 1e3, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "u8", "data": "synthetic", "config": {"workload": "copy 256 MiB device to device",
                                                           "bytes_read": nbytes, "bytes_written": nbytes},
            "roofline": roofline(2 * nbytes, avg)}


def copy_ceiling():
    """Practical streaming ceiling of this GPU: a 1 GiB device-to-device copy
    (past the last-level cache) by the library's non-temporal vector copy
    kernel, GB/s of bytes read + written."""
    import cubismnova_amd as cnv
    from cubismnova_amd import _abi
    ev = HipEvents()
    nbytes = 1 << 30
    a, b = cnv.DeviceBuffer(nbytes), cnv.DeviceBuffer(nbytes)
    a.fill(np.float32, 1.0)
    st = cnv.Stream()
    for _ in range(2):
        _abi.call("cnv_memcpy_stream", b.ptr, a.ptr, nbytes, st.handle)
    st.synchronize()
    e0, e1 = ev.new(), ev.new()
    ev.record(e0, st)
    for _ in range(5):
        _abi.call("cnv_memcpy_stream", b.ptr, a.ptr, nbytes, st.handle)
    ev.record(e1, st)
    return 2 * nbytes / (ev.elapsed_ms(e0, e1) / 1e3 / 5) / 1e9


def fanout_ceiling(fanout=6, nbytes=512 << 20):
    """Streaming ceiling of a one-read, `fanout`-write pattern (the WENO5
    operator's: one input component, six state components): GB/s of bytes
    read + written by the library's non-temporal fanout copy kernel, 512 MiB
    read (past the last-level cache)."""
    import cubismnova_amd as cnv
    from cubismnova_amd import _abi
    ev = HipEvents()
    a, b = cnv.DeviceBuffer(nbytes), cnv.DeviceBuffer(nbytes * fanout)
    a.fill(np.float32, 1.0)
    st = cnv.Stream()
    for _ in range(2):
        _abi.call("cnv_memcpy_fanout", b.ptr, nbytes, a.ptr, nbytes, fanout, st.handle)
    st.synchronize()
    e0, e1 = ev.new(), ev.new()
    ev.record(e0, st)
    for _ in range(5):
        _abi.call("cnv_memcpy_fanout", b.ptr, nbytes, a.ptr, nbytes, fanout, st.handle)
    ev.record(e1, st)
    return (1 + fanout) * nbytes / (ev.elapsed_ms(e0, e1) / 1e3 / 5) / 1e9


def result_check(out):
    """The tolerances each workload's line is quoted at: FP32 vs FP64 rel-L2
    <= 1e-6 (diffusion after all steps, WENO5 on the worst state); the
    refined Jacobi solve at 1e-12 and equal to FP64 Jacobi to 1e-9; the
    batch equal to its single-grid entry points (diffusion bit for bit,
    WENO5 within 1e-6).  Returns a message for the first one missed, or ""."""
    if "fp32_rel_l2_vs_fp64" in out and not out["fp32_rel_l2_vs_fp64"] <= 1e-6:
        return "diffusion FP32 vs FP64 rel-L2 %g" % out["fp32_rel_l2_vs_fp64"]
    if "fp32_rel_l2_vs_fp64_worst_state" in out and not out["fp32_rel_l2_vs_fp64_worst_state"] <= 1e-6:
        return "WENO5 FP32 vs FP64 rel-L2 %g" % out["fp32_rel_l2_vs_fp64_worst_state"]
    if "batch_vs_single_grid_max_rel" in out and not out["batch_vs_single_grid_max_rel"] <= 1e-6:
        return "batch vs single-grid max rel %g" % out["batch_vs_single_grid_max_rel"]
    cv = out.get("convergence_32cube")
    if cv and not (cv["refined_rel_residual"] <= 1e-12 and cv["refined_vs_fp64_rel_l2"] <= 1e-9):
        return "refined Jacobi residual %g, vs FP64 %g" % (cv["refined_rel_residual"], cv["refined_vs_fp64_rel_l2"])
    for name, rec in out.get("workloads", {}).items():
        bad = result_check(rec)
        if bad:
            return name + ": " + bad
    return ""


# steps and warmup of the workloads that run beside the headline diffusion
SUB_WORKLOADS = (("weno5", 5, 1), ("jacobi", 3, 1), ("batch", 20, 2))


def line_fields(rec, steps, warmup):
    """The fields every printed line carries, for a workload run on its own."""
    out = {"metric": rec["metric"], "value": rec["value"], "unit": rec["unit"], "n_gpus": 1, "steps": steps,
           "warmup": warmup, "ms_per_step": rec["ms_per_step"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None}
    out.update({k: v for k, v in rec.items() if k not in out})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="all", choices=["all", "diffusion", "weno5", "batch", "jacobi", "copy"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--cpu-seconds", type=float, default=3.0,
                    help="budget of each CPU baseline sample run (at least 10 samples are taken)")
    ap.add_argument("--no-autotune", action="store_true", help="use the built-in tile table")
    ap.add_argument("--batch-form", default="auto", choices=["auto", "one", "classes", "serial"],
                    help="batch workload: the autotuner's form, or pin the single launch, the class launches "
                         "(overlapping) or the class launches one after another")
    tile = lambda v: tuple(int(x) for x in v.split(","))
    ap.add_argument("--tile-f32", type=tile, default=None, help="WG,ZCHUNK for the FP32 kernel of the diffusion or WENO5 workload, WG,ZCHUNK[,PREFETCH] for the Jacobi sweep (skips autotune)")
    ap.add_argument("--tile2-f32", type=tile, default=None,
                    help="WG,ZCHUNK for the FP32 two-step diffusion kernel (skips its autotune)")
    ap.add_argument("--tile2-f64", type=tile, default=None,
                    help="WG,ZCHUNK for the FP64 two-step diffusion kernel (skips its autotune)")
    ap.add_argument("--tile-f64", type=tile, default=None, help="WG,ZCHUNK for the FP64 kernel of the diffusion or WENO5 workload, WG,ZCHUNK[,PREFETCH] for the Jacobi sweep (skips autotune)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the library has no multi-GPU path (got --gpus %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.steps < 1:
        sys.exit("--steps must be >= 1")
    fns = {"weno5": bench_weno5, "batch": bench_batch, "jacobi": bench_jacobi}
    if args.workload == "copy":
        out = bench_copy(args)
    elif args.workload in fns:
        out = line_fields(fns[args.workload](args, args.steps, args.warmup), args.steps, args.warmup)
    else:
        out = bench_diffusion(args)
        print("diffusion done: %.1f GCUPS FP32" % out["value"], file=sys.stderr, flush=True)
        if args.workload == "all":
            out["workloads"] = {}
            for name, k, w in SUB_WORKLOADS:
                out["workloads"][name] = fns[name](args, k, w)
                print("%s done: %.1f GCUPS FP32" % (name, out["workloads"][name]["value"]), file=sys.stderr,
                      flush=True)
    bad = result_check(out)
    if bad:
        # a fast line over wrong results is not a measurement
        print(json.dumps(out), file=sys.stderr)
        sys.exit("bench.py: results outside the stated tolerance: " + bad)
    if args.workload != "copy":
        c = copy_ceiling()
        out["hbm_ceiling"] = {"copy_nt_1GiB_GB_s": round(c, 1), "frac_of_peak": round(c / HBM_PEAK_GBS, 4),
                              "kernel_frac_of_copy": round(out["roofline"]["achieved"] / c, 4),
                              "copy": "1 GiB device-to-device, non-temporal 16-byte loads and stores "
                                      "(cnv_memcpy_stream)"}
        for rec in out.get("workloads", {}).values():
            rec["roofline"]["frac_of_copy"] = round(rec["roofline"]["achieved"] / c, 4)
        w = out["workloads"].get("weno5") if "workloads" in out else (out if args.workload == "weno5" else None)
        if w is not None:
            f = fanout_ceiling()
            w["roofline"]["fanout6_ceiling_GB_s"] = round(f, 1)
            w["roofline"]["frac_of_fanout6"] = round(w["roofline"]["achieved"] / f, 4)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
