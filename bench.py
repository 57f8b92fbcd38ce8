#!/usr/bin/env python3
"""RAAR(+ER) start-iterations per second on one MI355X, batched reduced-precision storage
against the plain fp32 one-start-per-launch path, with the HBM roofline of the dominant
kernel and the CPU oracle timed on the host beside it.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--configs c2,c4] [--no-cpu]

A step is one full recipe run over every start of the workload (the reference's
bench/bench.cpp:33-45 protocol: construct, warm up with a run of the same length so the
hipGraphs exist, then time).  The headline value is the 256^2 workload (32 starts,
repeat(10, RAAR(50, beta 0.9), Shrinkwrap) then ER(200), Poisson data with a beamstop)
on the batched fp16-storage path; the 128^3 workload (8 starts, repeat(8, RAAR(50),
Shrinkwrap rms) then ER(200)) is reported beside it under "c4".  Prints one compact JSON
line on stdout (the headline, the roofline of the dominant kernel and the CPU baseline);
the full per-workload record (every run, pass timing, per-start errors and the autotune
log) goes to the file named by --details, and a readable summary goes to stderr.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_HBM_GBS = 8000.0  # MI355X HBM3E, public spec

# Algorithmic bytes per element per start of each pass (the two-pass model): reads and
# writes of the wave, the iterate, the amplitude and the support.
PASS_BYTES = {
    ("modulus", "f32"): 8 + 4 + 8,
    ("update", "f32"): 8 + 8 + 1 + 8 + 8,
    ("inv", "f32"): 16,
    ("fwd", "f32"): 16,
    ("modulus", "f16"): 4 + 4 + 4,
    ("update", "f16"): 4 + 4 + 1 + 4 + 4,
    ("inv", "f16"): 8,
    ("fwd", "f16"): 8,
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name):
    from qoraphasecu_amd import workloads

    return workloads.config(name)


def time_ensemble(w, variant, steps, warmup):
    """All problems in one MultiSolver: every shape advances in the same graph launch."""
    from qoraphasecu_amd import _core

    ops = w["recipe"].ops()
    iters = w["recipe"].iterations()
    inten = [p[1] for p in w["problems"]]
    masks = [p[2] for p in w["problems"]]
    t0 = time.perf_counter()
    s = _core.MultiSolver(inten, None, masks, w["starts"], 0, 0, 0.1, "hip", variant, 10)
    build = time.perf_counter() - t0
    for _ in range(max(1, warmup)):
        s.run(ops)
    t0 = time.perf_counter()
    for _ in range(steps):
        s.run(ops)
    wall = time.perf_counter() - t0
    runs = w["starts"] * len(inten)
    rate = runs * iters * steps / wall
    best = [min(float(r[2][-1]) for r in per) for per in s.results()]
    log(f"  c3 ensemble {variant:>8}: {wall / steps * 1e3:9.2f} ms/run  {rate:12.4g} start-it/s  "
        f"(construction + tuning {build:.2f} s; median best error {np.median(best):.3e})")
    return dict(ms_per_run=wall / steps * 1e3, start_iterations_per_s=rate, construction_s=build,
                best_final_errors=best)


def time_per_problem(w, variant, steps):
    """The problems one after another, each batched over its own starts."""
    from qoraphasecu_amd import _core

    ops = w["recipe"].ops()
    iters = w["recipe"].iterations()
    solvers = [_core.Solver(p[1], None, p[2], w["starts"], 0, 0, None, 0.1, "hip", variant, 10)
               for p in w["problems"]]
    for s in solvers:
        s.run(ops)
    t0 = time.perf_counter()
    for _ in range(steps):
        for s in solvers:
            s.run(ops)
    wall = time.perf_counter() - t0
    rate = w["starts"] * len(solvers) * iters * steps / wall
    log(f"  c3 per-problem {variant:>8}: {wall / steps * 1e3:9.2f} ms/run  {rate:12.4g} start-it/s")
    return dict(ms_per_run=wall / steps * 1e3, start_iterations_per_s=rate)


def time_refinement(steps=500):
    """Config 5: Adam (lr 1e-2) on the support-projected iterate of 16 objects of 256^2,
    started from 50 RAAR iterations; the amplitude loss and its adjoint on the GPU."""
    import torch

    import qoraphasecu_amd as qpc
    from qoraphasecu_amd import autograd, workloads

    x0, sup, inten, masks = [], [], [], []
    for seed in range(16):
        obj, i, m = workloads.problem_2d(256, seed=seed)
        r = qpc.reconstruct(i, mask=m.astype(bool), recipe=qpc.Recipe(qpc.RAAR(50)), backend="hip", variant="batched")
        x0.append(r.object)
        sup.append(r.support)
        inten.append(i)
        masks.append(m.astype(bool))
    x0 = np.stack(x0)
    res = {}
    for graph in (True, False):
        autograd.refine(x0, np.stack(inten), np.stack(sup), mask=np.stack(masks), steps=5, graph=graph)  # warm up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out, hist = autograd.refine(x0, np.stack(inten), np.stack(sup), mask=np.stack(masks), steps=steps, graph=graph)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        h = hist.cpu().numpy()
        tag = "graph" if graph else "eager"
        log(f"  c5 adam ({tag}): {wall / steps * 1e3:.3f} ms/step (16 x 256^2), loss {h[0].sum():.4g} -> {h[-1].sum():.4g}")
        res[tag] = dict(ms_per_step=wall / steps * 1e3, steps=steps, loss_first=float(h[0].sum()),
                        loss_last=float(h[-1].sum()))
    return res


def time_variant(w, variant, batch, steps, warmup):
    from qoraphasecu_amd import _core

    ops = w["recipe"].ops()
    iters = w["recipe"].iterations()
    s = _core.Solver(w["intensity"], None, w["mask"], w["starts"], batch, 0, None, 0.1, "hip", variant, 10)
    t0 = time.perf_counter()
    s.run(ops)  # first run: graph capture included
    first = time.perf_counter() - t0
    results = s.results()
    final = [float(r[2][-1]) for r in results]
    obj_err = None
    if w["name"] in ("c1", "c4"):  # noiseless |F object|^2: the true object is the answer
        import qoraphasecu_amd as qpc

        best = int(np.argmin(final))
        obj_err = float(qpc.aligned_error(results[best][0], w["object"]))
    for _ in range(max(0, warmup - 1)):
        s.run(ops)
    t0 = time.perf_counter()
    for _ in range(steps):
        s.run(ops)  # run() returns after the solver's stream has drained
    wall = time.perf_counter() - t0
    ms = wall / steps * 1e3
    rate = w["starts"] * iters * steps / wall
    extra = f", aligned_error of its object {obj_err:.3e}" if obj_err is not None else ""
    log(f"  {w['name']} {variant:>8} batch={batch}: {ms:9.2f} ms/run  {rate:12.4g} start-it/s  "
        f"(first run {first * 1e3:.1f} ms, best final error {min(final):.3e}{extra})")
    return dict(ms_per_run=ms, start_iterations_per_s=rate, first_run_ms=first * 1e3,
                best_final_error=min(final), final_errors=final, best_object_aligned_error=obj_err)


def time_plain_radix8(w, steps):
    """The plain fp32 path (one start per launch, graph-replayed) at radix 8: faster than the
    default radix 16 for one-start launches of 256-point lines, but a different rounding,
    so the API keeps radix 16 for every variant (a start's result must not depend on the
    launch split).  RAAR(0.9) iterations of the recipe's count, without Shrinkwrap."""
    from qoraphasecu_amd import _core, _qpch

    amp, mask, sup = _core.make_problem(w["intensity"], None, w["mask"], 0.1)
    f = np.fft.ifftshift
    iters = w["recipe"].iterations() // 10 * 10
    s = _qpch.Solver(f(amp).copy(), f(mask).copy(), f(sup).copy(), starts=w["starts"], batch=1, radix=8,
                     schedule="graph")

    def run():
        s.begin_run(iters)
        for _ in range(iters // 10):
            s.iterate(10, 0.9, -0.8, False)
        s.synchronize()

    run()
    t0 = time.perf_counter()
    for _ in range(steps):
        run()
    wall = (time.perf_counter() - t0) / steps
    s.close()
    rate = w["starts"] * iters / wall
    log(f"  {w['name']} plain fp32 radix 8: {wall * 1e3:9.2f} ms/run  {rate:12.4g} start-it/s  (RAAR only)")
    return dict(ms_per_run=wall * 1e3, start_iterations_per_s=rate, iterations=iters)


def persistent_roofline(w, storage, length=50, reps=10):
    """The production path's kernel at 2D: one persistent launch (twopass_iterations) runs
    both passes of `length` iterations (a recipe stage, RAAR(50)).  Its average duration
    is timed with HIP events on the solver's stream around `reps` graph-replayed launches
    (qpch_solver_time_iterations); bytes are the two-pass model's per iteration."""
    from qoraphasecu_amd import _core, _qpch

    amp, mask, sup = _core.make_problem(w["intensity"], None, w["mask"], 0.1)
    f = np.fft.ifftshift
    s = _qpch.Solver(f(amp).copy(), f(mask).copy(), f(sup).copy(), starts=w["starts"], batch=0,
                     storage=storage, mixed=storage != "f32", schedule="graph", lanes=0, tuned=True)
    low = storage != "f32"
    us_it = s.time_iterations(low=low, length=length, reps=reps)
    tuning = s.tuning
    s.close()
    per_launch = tuning["batch"] if tuning and tuning["batch"] > 0 else w["starts"]
    groups = w["starts"] // per_launch
    key = "f16" if low else "f32"
    n = int(np.prod(w["shape"]))
    b = (PASS_BYTES[("modulus", key)] + PASS_BYTES[("update", key)]) * n * per_launch * length
    launch_us = us_it * length / groups
    return dict(iteration_us=us_it, launch_us=launch_us, length=length, starts_per_launch=per_launch,
                algorithmic_bytes=b, achieved_gbs=b / (launch_us * 1e-6) / 1e9, launch=tuning)


def pass_roofline(w, storage, reps=50):
    """Average duration of each pass of one batched iteration, HIP events on the solver
    stream, and the iteration of the launch-per-pass schedule replayed from its graphs."""
    from qoraphasecu_amd import _core, _qpch

    amp, mask, sup = _core.make_problem(w["intensity"], None, w["mask"], 0.1)
    f = np.fft.ifftshift
    s = _qpch.Solver(f(amp).copy(), f(mask).copy(), f(sup).copy(), starts=w["starts"], batch=0,
                     storage=storage, mixed=storage != "f32", schedule="graph", lanes=0, tuned=True, persist=2)
    low = storage != "f32"
    passes = s.time_passes(low=low, reps=reps)
    # the same iteration replayed from the production hipGraphs (chunks of 10)
    graph_us = s.time_iterations(low=low, length=10, reps=20)
    tuning = s.tuning
    s.close()
    n = int(np.prod(w["shape"]))
    key = "f16" if low else "f32"
    out = []
    per_launch = tuning["batch"] if tuning and tuning["batch"] > 0 else w["starts"]
    for axis, mode, ms in passes:
        b = PASS_BYTES[(mode, key)] * n * per_launch
        out.append(dict(axis=axis, mode=mode, ms=ms, algorithmic_bytes=b, achieved_gbs=b / (ms * 1e-3) / 1e9,
                        launch=tuning, graph_iteration_us=graph_us))
    return out


def cpu_baseline(w, threads, seconds=10.0, starts=None):
    """The oracle (float, OpenMP) on a bounded sample of the workload: RAAR iterations on
    every start (or on `starts` of them), as many as fill about `seconds` of CPU time (at
    least the reference's max(5, iters/20), bench/bench.cpp:148), sized from one timed
    warm-up iteration."""
    import oracle

    oracle.set_threads(threads)
    amp, mask, sup = oracle.make_problem(w["intensity"], None, w["mask"])
    starts = w["starts"] if starts is None else starts
    solver = oracle.Solver(amp, mask, sup, starts=starts, seed=0)
    t0 = time.perf_counter()
    solver.run([("raar", 1, 0.9, 0.0, 0.0, "max", 0)])
    one = time.perf_counter() - t0
    iters = max(5, w["recipe"].iterations() // 20, int(seconds / max(one, 1e-6)))
    t0 = time.perf_counter()
    solver.run([("raar", iters, 0.9, 0.0, 0.0, "max", 0)])
    wall = time.perf_counter() - t0
    return dict(value=starts * iters / wall, starts=starts, iterations=iters, seconds=wall, threads=threads)


def numpy_fft_pair(shape, reps=20):
    x = (np.random.default_rng(0).standard_normal(shape) + 0j).astype(np.complex64)
    np.fft.ifftn(np.fft.fftn(x))
    t0 = time.perf_counter()
    for _ in range(reps):
        np.fft.ifftn(np.fft.fftn(x))
    return (time.perf_counter() - t0) / reps * 1e3


def input_hashes(w):
    """SHA-256 of the generated inputs (float32 intensity, uint8 mask), as the committed
    manifest tests/golden/workloads_manifest.json records them."""
    import hashlib

    def h(a):
        return None if a is None else hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()

    if "problems" in w:
        return {"intensity": [h(np.asarray(p[1], np.float32)) for p in w["problems"]],
                "mask": [h(np.asarray(p[2], np.uint8)) for p in w["problems"]]}
    return {"intensity": h(np.asarray(w["intensity"], np.float32)),
            "mask": h(None if w["mask"] is None else np.asarray(w["mask"], np.uint8))}


def rocfft_floor(shape, starts, reps=50):
    """A forward + inverse complex64 transform of every start (torch.fft, i.e. rocFFT),
    captured in a graph and replayed: the floor of one iteration's transforms, timed on
    the capturing stream with HIP events."""
    import torch

    x = torch.randn((starts, *shape), dtype=torch.complex64, device="cuda")
    dims = tuple(range(1, len(shape) + 1))
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for _ in range(3):
            y = torch.fft.ifftn(torch.fft.fftn(x, dim=dims), dim=dims)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    mode = "graph"
    try:
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            y = torch.fft.ifftn(torch.fft.fftn(x, dim=dims), dim=dims)
        run = g.replay
    except Exception as e:  # timing floor only: fall back to eager launches
        mode = f"eager ({type(e).__name__})"

        def run():
            torch.fft.ifftn(torch.fft.fftn(x, dim=dims), dim=dims)
    run()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        run()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    return dict(ms_per_pair=ms, starts=starts, mode=mode, start_pairs_per_s=starts / (ms * 1e-3))


def cpu_full_run(w, threads):
    """The oracle (float, OpenMP) running the whole recipe on every start, after a warm-up
    run of the same recipe (bench/bench.cpp:33-45)."""
    import oracle

    oracle.set_threads(threads)
    amp, mask, sup = oracle.make_problem(w["intensity"], None, w["mask"])
    ops = w["recipe"].ops()
    iters = w["recipe"].iterations()
    warm = oracle.Solver(amp, mask, sup, starts=w["starts"], seed=0)
    warm.run(ops)
    solver = oracle.Solver(amp, mask, sup, starts=w["starts"], seed=0)
    t0 = time.perf_counter()
    solver.run(ops)
    wall = time.perf_counter() - t0
    best = min(float(r[2][-1]) for r in solver.results())
    return dict(value=w["starts"] * iters / wall, seconds=wall, threads=threads, iterations=iters,
                starts=w["starts"], best_final_error=best)


def time_next_rows(steps):
    """The rows after the hot path: a non-power-of-two problem (240^2 = 16 x 15, 8 starts,
    RAAR(50)) through the smooth-length kernels, with the oracle's rate on the same
    problem beside it; a prime row length (64 x 1021, 8 starts, RAAR(50): the chirp-z
    passes); 448^2 and 360^2 (8 starts, RAAR(50): the smooth kernels' prime short form and
    even split); and the device make_problem (autocorrelation support) and aligned_error at
    256^2."""
    import qoraphasecu_amd as qpc
    from qoraphasecu_amd import _core, workloads

    obj, inten, mask = workloads.problem_2d(240, seed=1000, photons=1e7, beamstop_radius=6, box=90)
    ops = qpc.Recipe(qpc.RAAR(50)).ops()
    res = {"workload": "240^2, 8 starts, RAAR(50), Poisson 1e7 photons, 6 px beamstop"}
    for variant in ("graph", "batched"):
        s = _core.Solver(inten, None, mask, 8, 1 if variant == "graph" else 0, 0, None, 0.1, "hip", variant, 10)
        s.run(ops)
        t0 = time.perf_counter()
        for _ in range(steps):
            s.run(ops)
        wall = (time.perf_counter() - t0) / steps
        res[variant] = dict(ms_per_run=wall * 1e3, start_iterations_per_s=8 * 50 / wall)
    log(f"  240^2 (smooth length 16 x 15, line in registers): graph {res['graph']['start_iterations_per_s']:.4g} start-it/s, "
        f"batched {res['batched']['start_iterations_per_s']:.4g}")
    # a prime row axis (1021: Bluestein's transform in the generic passes)
    obj3, inten3, mask3 = workloads.problem_2d(64, seed=1001, photons=1e7, beamstop_radius=4, box=24)
    inten3 = np.pad(inten3, ((0, 0), (478, 479)))  # 64 x 1021, the measured spectrum in the middle
    mask3 = np.pad(mask3, ((0, 0), (478, 479)))
    s = _core.Solver(inten3, None, mask3, 8, 0, 0, None, 0.1, "hip", "batched", 10)
    s.run(ops)
    t0 = time.perf_counter()
    for _ in range(steps):
        s.run(ops)
    wall = (time.perf_counter() - t0) / steps
    res["prime_64x1021"] = dict(ms_per_run=wall * 1e3, start_iterations_per_s=8 * 50 / wall)
    log(f"  64 x 1021 (Bluestein rows), 8 starts, RAAR(50): {res['prime_64x1021']['start_iterations_per_s']:.4g} "
        f"start-it/s")
    # smooth lengths with a prime factor of 7 (448 = 64 x 7) and with an even split whose
    # first factor is not a power of two (360 = 24 x 15), 8 starts, RAAR(50), all starts in
    # one launch
    for n, key, label in ((448, "smooth_448", "448^2 (64 x 7)"), (360, "smooth_360", "360^2 (24 x 15)")):
        objn, intenn, maskn = workloads.problem_2d(n, seed=1002, photons=1e7, beamstop_radius=6, box=n * 3 // 8)
        s = _core.Solver(intenn, None, maskn, 8, 0, 0, None, 0.1, "hip", "batched", 10)
        s.run(ops)
        t0 = time.perf_counter()
        for _ in range(steps):
            s.run(ops)
        wall = (time.perf_counter() - t0) / steps
        res[key] = dict(ms_per_run=wall * 1e3, start_iterations_per_s=8 * 50 / wall)
        log(f"  {label}, 8 starts, RAAR(50): {res[key]['start_iterations_per_s']:.4g} start-it/s")
    obj2, inten2, mask2 = workloads.problem_2d(256, seed=1000, photons=1e7, beamstop_radius=6)
    _core.make_problem(inten2, None, mask2, 0.1)
    t0 = time.perf_counter()
    for _ in range(20):
        _core.make_problem(inten2, None, mask2, 0.1)
    res["make_problem_256_ms"] = (time.perf_counter() - t0) / 20 * 1e3
    a = obj2 * np.exp(0.3j)
    qpc.aligned_error(a, obj2)
    t0 = time.perf_counter()
    for _ in range(20):
        qpc.aligned_error(a, obj2)
    res["aligned_error_256_ms"] = (time.perf_counter() - t0) / 20 * 1e3
    log(f"  make_problem 256^2: {res['make_problem_256_ms']:.2f} ms; aligned_error 256^2: "
        f"{res['aligned_error_256_ms']:.2f} ms")
    return res


def cpu_next_rows(threads):
    """The oracle on the 240^2 problem of time_next_rows (RAAR(5) on all 8 starts)."""
    import oracle
    from qoraphasecu_amd import workloads

    oracle.set_threads(threads)
    obj, inten, mask = workloads.problem_2d(240, seed=1000, photons=1e7, beamstop_radius=6, box=90)
    amp, m, sup = oracle.make_problem(inten, None, mask)
    solver = oracle.Solver(amp, m, sup, starts=8, seed=0)
    solver.run([("raar", 1, 0.9, 0.0, 0.0, "max", 0)])
    t0 = time.perf_counter()
    solver.run([("raar", 5, 0.9, 0.0, 0.0, "max", 0)])
    wall = time.perf_counter() - t0
    return dict(value=8 * 5 / wall, threads=threads)


def traffic_for(kernel_key):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, when present."""
    path = os.path.join(HERE, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        table = json.load(fh)
    v = table.get(kernel_key)
    return None if v is None else v.get("hbm_bytes_per_launch")


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--configs", default="c2,c4,c3,c5,c1,next")
    ap.add_argument("--variants", default="graph,mixed,mixed_bf16")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-passes", action="store_true", help="skip the per-pass timing")
    ap.add_argument("--no-floor", action="store_true", help="skip the rocFFT transform floor")
    ap.add_argument("--details", default=None, metavar="PATH",
                    help="write the full per-workload record as JSON to PATH")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB 10 times (the reference point of the "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch

        a = torch.ones(args.calibrate_copy * (1 << 20) // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(10):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": a.numel() * 4, "launches": 10}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (starts shard across GPUs "
              "with no per-iteration exchange). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import qoraphasecu_amd as qpc

    info = qpc.hip_device_info()
    log(f"device: {info}")
    details = {"device": info}
    headline = None
    roof = None
    for name in [c for c in args.configs.split(",") if c]:
        if name == "c5":
            details["c5"] = time_refinement()
            continue
        if name == "next":
            res = time_next_rows(args.steps)
            if not args.no_cpu:
                res["cpu_oracle_240"] = cpu_next_rows(min(16, os.cpu_count() or 1))
                log(f"  240^2 cpu oracle: {res['cpu_oracle_240']['value']:.4g} start-it/s")
            details["next"] = res
            continue
        w = workload(name)
        if name == "c1":
            # the reference's CPU-runnable case: the oracle in full (all threads and one), and
            # the device on the same problem
            res = {"workload": w["desc"], "iterations": w["recipe"].iterations(), "starts": w["starts"],
                   "inputs_sha256": input_hashes(w)}
            for variant in ("graph", "batched"):
                res[variant] = time_variant(w, variant, 1 if variant == "graph" else 0, args.steps, args.warmup)
            if not args.no_cpu:
                threads = min(16, os.cpu_count() or 1)
                res["cpu_oracle"] = cpu_full_run(w, threads)
                res["cpu_oracle_1_thread"] = cpu_full_run(w, 1)
                log(f"  c1 cpu oracle: {res['cpu_oracle']['value']:.4g} start-it/s ({threads} threads), "
                    f"{res['cpu_oracle_1_thread']['value']:.4g} (1 thread); best final error "
                    f"{res['cpu_oracle']['best_final_error']:.3e}")
            details["c1"] = res
            continue
        if name == "c3":
            res = {"workload": w["desc"], "iterations": w["recipe"].iterations(), "starts": w["starts"],
                   "problems": len(w["problems"]), "inputs_sha256": input_hashes(w)}
            res["ensemble_mixed"] = time_ensemble(w, "mixed", args.steps, args.warmup)
            res["ensemble_batched"] = time_ensemble(w, "batched", args.steps, args.warmup)
            res["per_problem_batched"] = time_per_problem(w, "batched", args.steps)
            res["speedup_ensemble_mixed_vs_per_problem_fp32"] = (
                res["ensemble_mixed"]["start_iterations_per_s"] / res["per_problem_batched"]["start_iterations_per_s"])
            from qoraphasecu_amd import _qpch

            res["autotune_log"] = _qpch.autotune_log()
            details["c3"] = res
            continue
        log(f"{name}: {w['desc']}")
        res = {"workload": w["desc"], "iterations": w["recipe"].iterations(), "starts": w["starts"],
               "inputs_sha256": input_hashes(w)}
        for variant in [v for v in args.variants.split(",") if v]:
            batch = 1 if variant in ("graph", "fused", "naive", "callback") else 0
            res[variant] = time_variant(w, variant, batch, args.steps, args.warmup)
        if "graph" in res:
            for v in ("mixed", "mixed_bf16", "half", "bf16", "batched"):
                if v in res:
                    res[f"speedup_{v}_vs_plain"] = res[v]["start_iterations_per_s"] / res["graph"]["start_iterations_per_s"]
            if len(w["shape"]) == 2:
                res["graph_radix8"] = time_plain_radix8(w, args.steps)
                if "mixed" in res:
                    res["speedup_mixed_vs_plain_radix8"] = (res["mixed"]["start_iterations_per_s"]
                                                           / res["graph_radix8"]["start_iterations_per_s"])
        n = int(np.prod(w["shape"]))
        for v in ("mixed", "mixed_bf16", "batched", "graph"):
            if v not in res:
                continue
            # whole-run HBM model: RAAR iterations at the storage's bytes, ER at 53 B
            iters_total = w["recipe"].iterations()
            er = 200
            low_b = 29 if v in ("mixed", "mixed_bf16", "half", "bf16") else 53
            model_bytes = n * w["starts"] * ((iters_total - er) * low_b + er * 53)
            gbs = model_bytes / (res[v]["ms_per_run"] * 1e-3) / 1e9
            res[v]["model_gbs"] = gbs
            res[v]["model_hbm_frac"] = gbs / PEAK_HBM_GBS
        passes16 = [] if args.no_passes else pass_roofline(w, "f16")
        passes32 = [] if args.no_passes else pass_roofline(w, "f32")
        res["passes_f16"] = passes16
        res["passes_f32"] = passes32
        persist16 = None
        if len(w["shape"]) == 2 and not args.no_passes:
            persist16 = persistent_roofline(w, "f16")
            res["persistent_f16"] = persist16
            res["persistent_f32"] = persistent_roofline(w, "f32")
            for tag in ("f16", "f32"):
                pr = res[f"persistent_{tag}"]
                log(f"    {tag} persistent launch ({pr['length']} iterations, {pr['starts_per_launch']} starts): "
                    f"{pr['launch_us']:8.1f} us  {pr['achieved_gbs']:8.0f} GB/s  ({pr['iteration_us']:.2f} us per iteration)")
        for tag, ps in (("f16", passes16), ("f32", passes32)):
            for p in ps:
                log(f"    {tag} pass axis {p['axis']} {p['mode']:>8}: {p['ms'] * 1e3:8.1f} us  {p['achieved_gbs']:8.0f} GB/s")
            if ps:
                per_launch = ps[0]["launch"]["batch"] if ps[0]["launch"] and ps[0]["launch"]["batch"] > 0 else w["starts"]
                eager = sum(p["ms"] for p in ps) * 1e3 * (w["starts"] // per_launch)
                log(f"    {tag} iteration of all starts: passes launched one by one {eager:.1f} us, "
                    f"replayed from the hipGraphs {ps[0]['graph_iteration_us']:.1f} us")
        if not args.no_floor:
            fl = rocfft_floor(w["shape"], w["starts"])
            res["rocfft_pair_floor"] = fl
            it_us = {v: res[v]["ms_per_run"] * 1e3 / w["recipe"].iterations() for v in ("graph", "mixed") if v in res}
            log(f"    rocFFT fwd+inv pair over all {w['starts']} starts ({fl['mode']}): {fl['ms_per_pair'] * 1e3:.1f} us; "
                + ", ".join(f"{v} iteration {t:.1f} us" for v, t in it_us.items()))
        if not args.no_cpu:
            threads = min(16, os.cpu_count() or 1)
            cb = cpu_baseline(w, threads)
            res["cpu_oracle"] = cb
            # one thread on one start, as the reference's single-thread CPU timing
            res["cpu_oracle_1thread"] = cpu_baseline(w, 1, seconds=3.0, starts=1)
            res["numpy_fft_pair_ms"] = numpy_fft_pair(w["shape"])
            log(f"    cpu oracle ({threads} threads): {cb['value']:.4g} start-it/s; 1 thread: "
                f"{res['cpu_oracle_1thread']['value']:.4g}; numpy fft pair {res['numpy_fft_pair_ms']:.2f} ms")
        details[name] = res
        if name == "c2":
            headline = res
        if name == "c2" and persist16 and persist16["launch"] and persist16["launch"].get("persist") == 2:
            # the tuner chose the launch-per-pass schedule on this device: its dominant pass
            longest = max(p["ms"] for p in passes16)
            dom = max((p for p in passes16 if p["ms"] >= 0.98 * longest), key=lambda p: p["algorithmic_bytes"])
            roof = {
                "bound": "hbm",
                "achieved": dom["achieved_gbs"],
                "peak": PEAK_HBM_GBS,
                "unit": "GB/s",
                "frac": dom["achieved_gbs"] / PEAK_HBM_GBS,
                "traffic": traffic_for(f"c2_f16_{dom['mode']}"),
                "kernel": f"line_pass {dom['mode']} (axis {dom['axis']}, fp16 storage, 32 starts per launch; "
                          "the tuner chose a launch per pass)",
                "algorithmic_bytes_per_launch": dom["algorithmic_bytes"],
                "avg_launch_ms": dom["ms"],
                "iteration_us": persist16["iteration_us"],
                "pass_launch_iteration_us": passes16[0]["graph_iteration_us"],
                "passes": [{"kernel": f"line_pass {p['mode']} (axis {p['axis']})", "us": p["ms"] * 1e3,
                            "frac": p["achieved_gbs"] / PEAK_HBM_GBS} for p in passes16],
            }
        elif name == "c2" and persist16:
            # The dominant kernel of the production path: the persistent launch that runs
            # every iteration of a RAAR(50) stage (91% of the recipe's iterations at fp16
            # storage).  The launch-per-pass kernels it replaces are reported beside it.
            roof = {
                "bound": "hbm",
                "achieved": persist16["achieved_gbs"],
                "peak": PEAK_HBM_GBS,
                "unit": "GB/s",
                "frac": persist16["achieved_gbs"] / PEAK_HBM_GBS,
                "traffic": traffic_for("c2_f16_twopass"),
                "kernel": f"twopass_iterations (both passes of {persist16['length']} iterations, fp16 storage, "
                          f"{persist16['starts_per_launch']} starts per launch)",
                "algorithmic_bytes_per_launch": persist16["algorithmic_bytes"],
                "avg_launch_ms": persist16["launch_us"] * 1e-3,
                "iteration_us": persist16["iteration_us"],
                "pass_launch_iteration_us": passes16[0]["graph_iteration_us"] if passes16 else None,
                "passes": [{"kernel": f"line_pass {p['mode']} (axis {p['axis']})", "us": p["ms"] * 1e3,
                            "frac": p["achieved_gbs"] / PEAK_HBM_GBS} for p in passes16],
            }

    from qoraphasecu_amd import _qpch

    details["autotune_log"] = _qpch.autotune_log().splitlines()
    if args.details:
        os.makedirs(os.path.dirname(os.path.abspath(args.details)), exist_ok=True)
        with open(args.details, "w") as fh:
            json.dump(details, fh, default=float)
    if headline is None:
        print(json.dumps({"metric": "RAAR iters/s (all starts) at 256² & 128³, batched/fp16 vs plain fp32; %HBM",
                          "value": None, "note": "the 256^2 workload (c2) was not run"}))
        return
    best = "mixed" if "mixed" in headline else [k for k in ("mixed_bf16", "batched", "graph") if k in headline][0]
    h = headline[best]
    cpu = headline.get("cpu_oracle")

    def rnd(v, k=4):
        return None if v is None else float(f"{v:.{k}g}")

    out = {
        "metric": "RAAR iters/s (all starts) at 256² & 128³, batched/fp16 vs plain fp32; %HBM",
        "value": rnd(h["start_iterations_per_s"], 6),
        "unit": "start-iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": rnd(h["ms_per_run"], 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "storage": "f16 (ER in f32)" if best == "mixed" else best,
        "data": "synthetic (seeded ellipse objects, Poisson counts)",
        "config": {
            "workload": "256x256, 32 starts, repeat(10, RAAR(50, 0.9), Shrinkwrap) + ER(200), Poisson 1e7 photons, "
                        f"6 px beamstop; {best}, all starts in one launch",
            "starts": headline["starts"],
            "iterations": headline["iterations"],
        },
        "plain_fp32": rnd(headline["graph"]["start_iterations_per_s"]) if "graph" in headline else None,
        "speedup_vs_plain_fp32": rnd(headline.get(f"speedup_{best}_vs_plain"), 3),
        "model_hbm_frac": rnd(h.get("model_hbm_frac"), 3),
        "roofline": None if roof is None else {
            "bound": roof["bound"],
            "achieved": rnd(roof["achieved"]),
            "peak": roof["peak"],
            "unit": roof["unit"],
            "frac": rnd(roof["frac"], 3),
            "traffic": roof["traffic"],
            "kernel": roof["kernel"],
            "bytes_per_launch": roof["algorithmic_bytes_per_launch"],
            "avg_launch_us": rnd(roof["avg_launch_ms"] * 1e3),
            "iteration_us": rnd(roof["iteration_us"]),
            "pass_launches": {"iteration_us": rnd(roof["pass_launch_iteration_us"]),
                              "passes": [{"kernel": q["kernel"], "us": rnd(q["us"], 3), "frac": rnd(q["frac"], 3)}
                                         for q in roof["passes"]]},
        },
        "cpu_baseline": None if cpu is None else {
            "value": rnd(cpu["value"]),
            "unit": "start-iterations/s",
            "cores": cpu["threads"],
            "kind": "port",
            "sample": f"oracle (C, float, OpenMP) RAAR x{cpu['iterations']} on all {cpu['starts']} starts "
                      f"of the 256^2 workload ({cpu['seconds']:.1f} s)",
        },
    }
    c4 = details.get("c4")
    if c4 is not None:
        best4 = "mixed" if "mixed" in c4 else None
        out["c4"] = {
            "workload": "128^3, 8 starts, repeat(8, RAAR(50), Shrinkwrap rms) + ER(200)",
            "mixed": rnd(c4["mixed"]["start_iterations_per_s"]) if best4 else None,
            "plain_fp32": rnd(c4["graph"]["start_iterations_per_s"]) if "graph" in c4 else None,
            "speedup": rnd(c4.get("speedup_mixed_vs_plain"), 3),
            "model_hbm_frac": rnd(c4["mixed"].get("model_hbm_frac"), 3) if best4 else None,
        }
    line = json.dumps(out, ensure_ascii=False)
    if len(line) > 2048:
        raise RuntimeError(f"bench line is {len(line)} bytes; keep it within 2 KB")
    print(line, flush=True)


if __name__ == "__main__":
    main()
