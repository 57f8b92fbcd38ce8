"""Recipe helpers shared by the tests: the 7-tuples of the reference's Python binding
(python/bindings.cpp:22-23) and a walker that drives the C-ABI solver the way the C++
HipSolver does (a mirror used to test the ABI on its own)."""

import numpy as np

import oracle


def er(n):
    return ("er", n, 0.0, 0.0, 0.0, "max", 0)


def hio(n, beta=0.9):
    return ("hio", n, beta, 0.0, 0.0, "max", 0)


def raar(n, beta=0.9):
    return ("raar", n, beta, 0.0, 0.0, "max", 0)


def shrinkwrap(threshold=0.3, sigma=1.0, method="max"):
    return ("shrinkwrap", 0, 0.0, threshold, sigma, method, 0)


def detwin(*ops, axis=0):
    return [("detwin_begin", 0, 0.0, 0.0, 0.0, "max", axis), *ops, ("detwin_end", 0, 0.0, 0.0, 0.0, "max", 0)]


def flat(*parts):
    out = []
    for p in parts:
        if isinstance(p, list):
            out.extend(p)
        else:
            out.append(p)
    return out


def repeat(n, *parts):
    return flat(*parts) * n


def default_recipe():
    """src/core/recipe.cpp:26-28"""
    return flat(raar(50), detwin(hio(20)), raar(50), repeat(8, raar(50), shrinkwrap()), er(200))


def iterations(ops):
    return sum(op[1] for op in ops if op[0] in ("er", "hio", "raar"))


def coefficients(kind, beta):
    """src/core/pointwise.hpp:41-45 in float32"""
    f = np.float32
    if kind == "er":
        return 0.0, 0.0
    if kind == "hio":
        return 1.0, float(-f(beta))
    return float(f(beta)), float(f(1.0) - f(2.0) * f(beta))


def run_abi(solver, ops, sync_every=10):
    """Walk a recipe over a qoraphasecu_amd._qpch.Solver; returns errors [iters, starts]."""
    total = iterations(ops)
    solver.begin_run(total)
    done = 0
    for op in ops:
        kind, n, beta, thr, sigma, method, axis = op
        if kind in ("er", "hio", "raar"):
            a, b = coefficients(kind, beta)
            k = 0
            while k < n:
                ln = min(n - k, sync_every - done % sync_every)
                solver.iterate(ln, a, b, refine=(kind == "er"))
                k += ln
                done += ln
        elif kind == "shrinkwrap":
            solver.shrinkwrap(thr, sigma, method)
        elif kind == "detwin_begin":
            solver.save_support()
            s = solver.get_support()
            for k in range(s.shape[0]):
                c = oracle.to_centered(s[k])
                s[k] = oracle.to_fft_order(oracle.half_support(c, axis))
            solver.set_support(s)
        elif kind == "detwin_end":
            solver.restore_support()
    num = solver.read_errors(0, total)
    return np.sqrt(num / (solver.energy + 1e-12)).astype(np.float32)
