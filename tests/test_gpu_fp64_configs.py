"""The benchmark workloads against an FP64 oracle run from the same starts (same seed, same
support and mask): the per-iteration Fourier error history, the final error and the object
error (aligned_error) of the fp32 one-start-per-launch path and the batched fp16 /
bf16-storage paths.  Each config runs its full recipe on a few starts (the FP64 oracle is
the slow side).  With QPCH_REPORT=<file> the numbers are written as JSON.

HIO / RAAR with Shrinkwrap and detwin is chaotic: from the same start, two runs that differ
only in rounding separate after some iterations and end in different local minima.  The
yardstick is float rounding itself, measured on the same starts by two independent float
runs followed against the oracle's double run: the oracle's own float run and, on the 2D
workloads, a numpy float32 run (pocketfft, the reference's CPU FFT library;
tests/numpy_float32.py).  The two differ by up to 8 iterations from one problem to the
next (256^2 ensemble problem: 39-42 against 34).  The device fp32 paths must follow the
FP64 error history (to 1e-3 relative) for at least 80% as many iterations as the shorter
of the two float references does, the fp16 / bf16-storage paths must agree with it over the
first iterations to their storage precision, and where the dynamics converge (the noiseless
128^3 volume) the final error must be at least as close to the FP64 run's as the CPU float
run's is, and is never required closer than a floor (1e-4 for fp32 storage, 1e-3 for fp16 /
bf16).  At 128^3 the CPU float run itself ends 4.6e-4 from the FP64 run (final errors
0.02919 and 0.02873)."""

import json
import os

import numpy as np
import pytest

import numpy_float32
import oracle

pytestmark = pytest.mark.gpu

# (config, starts, variants, floor of the final-error bar (absolute) where the run converges,
# or None)
CASES = [("c1", 8, ("graph", "batched", "mixed"), None),
         ("c2", 4, ("graph", "mixed", "mixed_bf16"), None),
         ("c4", 1, ("graph", "mixed", "mixed_bf16"), {"graph": 1e-4, "mixed": 1e-3, "mixed_bf16": 1e-3})]
LOW_TOL = 3e-2      # fp16 / bf16 storage: relative error difference over the first iterations
LOW_ITERS = 3
_REPORT = {}


@pytest.fixture(scope="module", autouse=True)
def _write_report():
    yield
    path = os.environ.get("QPCH_REPORT")
    if path and _REPORT:
        with open(path, "w") as fh:
            json.dump(_REPORT, fh, indent=1)


def _tracking(hist, ref, tol=1e-3):
    """Iterations before the relative difference of the two error histories first exceeds tol."""
    d = np.abs(np.asarray(hist, np.float64) - ref) / np.maximum(ref, 1e-30)
    bad = np.nonzero(d > tol)[0]
    return int(bad[0]) if bad.size else int(len(ref))


@pytest.mark.parametrize("name,starts,variants,final_bar", CASES)
def test_against_fp64_oracle(name, starts, variants, final_bar):
    import qoraphasecu_amd as qpc
    from qoraphasecu_amd import workloads

    w = workloads.config(name)
    oracle.set_threads(min(16, os.cpu_count() or 1))
    amp, mask, sup = oracle.make_problem(w["intensity"], None, w["mask"])
    ref = oracle.Solver(amp, mask, sup, starts=starts, seed=0, prec="d")
    ref.run(w["recipe"].ops())
    want = ref.results()
    ref_hist = [np.asarray(r[2], np.float64) for r in want]
    best_ref = min(float(h[-1]) for h in ref_hist)
    yard = oracle.Solver(amp, mask, sup, starts=starts, seed=0, prec="f")
    yard.run(w["recipe"].ops())
    yard_res = yard.results()
    yard_track = [_tracking(r[2], h) for r, h in zip(yard_res, ref_hist)]
    if len(w["shape"]) == 2:  # the second float reference (at 128^3 it would take minutes)
        np_runs = numpy_float32.run(amp, mask, sup, numpy_float32.random_starts(amp, sup, starts), w["recipe"].ops())
        np_track = [_tracking(h2, h) for h2, h in zip(np_runs, ref_hist)]
    else:
        np_track = yard_track
    yard_final = abs(min(float(r[2][-1]) for r in yard_res) - best_ref)
    rep = {"starts": starts, "iterations": w["recipe"].iterations(),
           "fp64_final_errors": [float(h[-1]) for h in ref_hist], "fp64_best_final_error": best_ref,
           "oracle_float_iterations_within_1e-3": yard_track, "numpy_float32_iterations_within_1e-3": np_track,
           "oracle_float_best_final_abs_diff": yard_final}
    print(f"{name}: the oracle's float run tracks its double run to 1e-3 for {min(yard_track)}-{max(yard_track)} "
          f"iterations, numpy float32 for {min(np_track)}-{max(np_track)}")
    for variant in variants:
        got = qpc.reconstruct(w["intensity"], support=sup.astype(bool),
                              mask=None if w["mask"] is None else w["mask"].astype(bool), recipe=w["recipe"],
                              starts=starts, seed=0, backend="hip", variant=variant)
        got = [got] if starts == 1 else list(got)
        errs = [float(g.error) for g in got]
        assert all(np.isfinite(errs)), (variant, errs)
        low = variant not in ("graph", "batched")
        early = max(float(np.max(np.abs(g.errors[:LOW_ITERS] - h[:LOW_ITERS]) / h[:LOW_ITERS]))
                    for g, h in zip(got, ref_hist))
        track = [_tracking(g.errors, h) for g, h in zip(got, ref_hist)]
        obj = [float(qpc.aligned_error(g.object, r[0].astype(np.complex64))) for g, r in zip(got, want)]
        best_gpu = min(errs)
        rep[variant] = {"final_errors": errs, "best_final_error": best_gpu, "best_abs_diff": abs(best_gpu - best_ref),
                        "per_start_abs_diff": [abs(e - float(h[-1])) for e, h in zip(errs, ref_hist)],
                        f"max_rel_diff_first_{LOW_ITERS}_iterations": early, "iterations_within_1e-3": track,
                        "object_error_vs_fp64": obj}
        print(f"{name} {variant}: best final error {best_gpu:.4e} (fp64 {best_ref:.4e}, |diff| "
              f"{abs(best_gpu - best_ref):.2e}); first {LOW_ITERS} iterations within {early:.1e}; tracks the fp64 "
              f"history to 1e-3 for {min(track)}-{max(track)} iterations; object error vs fp64 median "
              f"{float(np.median(obj)):.2e}")
        if low:
            assert early <= LOW_TOL, (variant, early)
        else:
            # the iterations to divergence grow with the log of the per-iteration rounding:
            # 80% of the CPU float reference's count allows a few times its rounding error
            assert min(track) >= 0.8 * min(min(yard_track), min(np_track)), (variant, track, yard_track, np_track)
        if final_bar is not None:
            bar = max(final_bar[variant], yard_final)
            assert abs(best_gpu - best_ref) <= bar, (variant, best_gpu, best_ref, yard_final)
    _REPORT[name] = rep


ENSEMBLE_SAMPLE = (0, 16, 32)   # the first problem of each size (128^2, 256^2, 512^2)
ENSEMBLE_STARTS = 2


def test_ensemble_against_fp64_oracle():
    """The uneven ensemble: a problem of each size, run together (one device solver per shape,
    one graph launch per chunk) on the fp32 and fp16-storage paths, each problem against its
    own FP64 oracle run from the same starts.  The recipe (the default one: RAAR, detwin,
    Shrinkwrap, ER) is chaotic, so the bars are those of the 64^2 and 256^2 workloads."""
    import qoraphasecu_amd as qpc
    from qoraphasecu_amd import workloads

    w = workloads.config("c3")
    oracle.set_threads(min(16, os.cpu_count() or 1))
    ops = w["recipe"].ops()
    picked = [w["problems"][k] for k in ENSEMBLE_SAMPLE]
    refs = []
    for _, intensity, mask in picked:
        amp, m, sup = oracle.make_problem(intensity, None, mask)
        ref = oracle.Solver(amp, m, sup, starts=ENSEMBLE_STARTS, seed=0, prec="d")
        ref.run(ops)
        want = ref.results()
        hist = [np.asarray(r[2], np.float64) for r in want]
        yard = oracle.Solver(amp, m, sup, starts=ENSEMBLE_STARTS, seed=0, prec="f")
        yard.run(ops)
        yard_track = [_tracking(r[2], h) for r, h in zip(yard.results(), hist)]
        np_runs = numpy_float32.run(amp, m, sup, numpy_float32.random_starts(amp, sup, ENSEMBLE_STARTS), ops)
        np_track = [_tracking(h2, h) for h2, h in zip(np_runs, hist)]
        refs.append(dict(sup=sup.astype(bool), want=want, hist=hist, yard_track=yard_track, np_track=np_track))
    rep = {"problems": [f"{p[1].shape[0]}^2 (problem {k})" for k, p in zip(ENSEMBLE_SAMPLE, picked)],
           "starts": ENSEMBLE_STARTS, "iterations": w["recipe"].iterations(),
           "fp64_best_final_errors": [min(float(h[-1]) for h in r["hist"]) for r in refs],
           "oracle_float_iterations_within_1e-3": [r["yard_track"] for r in refs],
           "numpy_float32_iterations_within_1e-3": [r["np_track"] for r in refs]}
    for variant in ("batched", "mixed"):
        runs = qpc.reconstruct_many([p[1] for p in picked], supports=[r["sup"] for r in refs],
                                    masks=[p[2].astype(bool) for p in picked], recipe=w["recipe"],
                                    starts=ENSEMBLE_STARTS, seed=0, backend="hip", variant=variant)
        rep[variant] = []
        for (k, p), run, r in zip(zip(ENSEMBLE_SAMPLE, picked), runs, refs):
            got = list(run)
            errs = [float(g.error) for g in got]
            assert all(np.isfinite(errs)), (variant, k, errs)
            early = max(float(np.max(np.abs(g.errors[:LOW_ITERS] - h[:LOW_ITERS]) / h[:LOW_ITERS]))
                        for g, h in zip(got, r["hist"]))
            track = [_tracking(g.errors, h) for g, h in zip(got, r["hist"])]
            obj = [float(qpc.aligned_error(g.object, o[0].astype(np.complex64))) for g, o in zip(got, r["want"])]
            best_ref = min(float(h[-1]) for h in r["hist"])
            rep[variant].append({"best_final_error": min(errs), "best_abs_diff": abs(min(errs) - best_ref),
                                 f"max_rel_diff_first_{LOW_ITERS}_iterations": early,
                                 "iterations_within_1e-3": track, "object_error_vs_fp64": obj})
            print(f"c3 problem {k} ({p[1].shape[0]}^2) {variant}: best final error {min(errs):.4e} (fp64 "
                  f"{best_ref:.4e}); first {LOW_ITERS} iterations within {early:.1e}; tracks the fp64 history to "
                  f"1e-3 for {min(track)}-{max(track)} iterations (oracle float run: {min(r['yard_track'])}-"
                  f"{max(r['yard_track'])}, numpy float32: {min(r['np_track'])}-{max(r['np_track'])}); object error "
                  f"vs fp64 median {float(np.median(obj)):.2e}")
            if variant == "mixed":
                assert early <= LOW_TOL, (variant, k, early)
            else:
                floor = min(min(r["yard_track"]), min(r["np_track"]))
                assert min(track) >= 0.8 * floor, (variant, k, track, r["yard_track"], r["np_track"])
    _REPORT["c3"] = rep
