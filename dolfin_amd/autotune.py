"""Autotuner for the blocked-scatter plans and the batched launch.

The tunables are the scatter strategy (owner-computes blocks, graph colouring, global
atomics), the plan granularity (P1 rows per block, P1 workgroup size, vector-P2 LDS values
per block -- i.e. cells per workgroup) and the launch mode of the batched kernels (one
workgroup per row block, or a number of persistent workgroups per CU). Candidates are timed with HIP events on a sample of the
workload and the fastest is returned; results are cached per (kind, size class).
"""
import ctypes
import itertools
import json
import os
import time

import torch

from . import _lib

_CACHE = {}

P1_ROWS = (128, 160, 192, 256)
P1_THREADS = (256, 512)
P2_VALS = (4096, 4608, 6144, 8192)


def set_plan_defaults(p1_rows=0, p1_threads=0, p2_vals=0):
    _lib.check(_lib.load().dfa_set_plan_defaults(int(p1_rows), int(p1_threads), int(p2_vals)))


def _time(fn, reps=5, warmup=2):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(s)
    for _ in range(reps):
        fn()
    b.record(s)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def tune_forms(make_forms, candidates=None, reps=5):
    """make_forms() -> list of forms built from fresh spaces (so plans are rebuilt with the
    current defaults). Each candidate (p1_rows, p1_threads, p2_vals) is applied with
    set_plan_defaults, the forms are assembled one by one and timed; returns
    (best candidate, {candidate: ms})."""
    from .fem import Assembler
    from .la import Matrix
    if candidates is None:
        candidates = list(itertools.product(P1_ROWS, P1_THREADS, P2_VALS))
    times = {}
    for cand in candidates:
        set_plan_defaults(*cand)
        try:
            forms = make_forms()
            tensors = [Matrix() for _ in forms]
            asm = Assembler()

            def run():
                for f, A in zip(forms, tensors):
                    asm.assemble(A, f)
            times[cand] = _time(run, reps)
        finally:
            set_plan_defaults(0, 0, 0)
    best = min(times, key=times.get)
    return best, times


def tune_scatter(forms, scatters=None, reps=5):
    """Time the assembly of `forms` (rank-2 forms, each into its own matrix) with every
    scatter strategy the element supports and return (fastest, {strategy: ms}). Plans are
    built by an untimed first call."""
    from .fem import Assembler
    from .la import Matrix
    if scatters is None:
        p1 = all(f.V.element == _lib.ELEMENT_P1 for f in forms)
        scatters = ((_lib.SCATTER_BLOCKED, _lib.SCATTER_COLORED, _lib.SCATTER_ATOMIC) if p1
                    else (_lib.SCATTER_BLOCKED, _lib.SCATTER_ATOMIC))
    times = {}
    for sc in scatters:
        tensors = [Matrix() for _ in forms]
        asm = Assembler(sc)

        def run():
            for f, A in zip(forms, tensors):
                asm.assemble(A, f)
        times[sc] = _time(run, reps)
    best = min(times, key=times.get)
    return best, times


def tune_batch_grid(make_batch, grids=(1, 2, 4, 0), reps=5):
    """Time the batched launch for several persistent-workgroups-per-CU counts (0: one
    workgroup per row block). make_batch(g) should reuse the members' spaces, whose plans are
    then already built."""
    times = {}
    for g in grids:
        b = make_batch(g)
        times[g] = _time(b.assemble, reps)
        del b
    best = min(times, key=times.get)
    return best, times


def tune_batch_plans(specs, make_form, p1_rows=(128, 160, 192), p2_vals=(4096, 4608, 6144),
                     reps=3):
    """Cells per workgroup of the batched launch: the row-block size of the P1 members and the
    LDS values per block of the vector-P2 members, timed through the batched launch itself on
    sample members (specs, built by make_form on fresh spaces so that every candidate builds
    its own plans). The two element kinds run as separate kernels of the launch, so each is
    tuned on its own members. Returns ((p1_rows, p2_vals), {"p1_rows": {rows: ms},
    "p2_vals": {vals: ms}})."""
    from .batch import BatchAssembler
    times = {"p1_rows": {}, "p2_vals": {}}
    for kind, key, cands in ((0, "p1_rows", p1_rows), (1, "p2_vals", p2_vals)):
        members = [sp for sp in specs if sp[0] == kind]
        if not members:
            continue
        for c in cands:
            b = BatchAssembler([make_form(sp) for sp in members],
                               plan_defaults=(c if kind == 0 else 0, 0, c if kind == 1 else 0))
            times[key][c] = _time(b.assemble, reps)
            del b
    best = tuple(min(times[k], key=times[k].get) if times[k] else 0 for k in ("p1_rows", "p2_vals"))
    return best, times


def cached(key, compute):
    if key not in _CACHE:
        _CACHE[key] = compute()
    return _CACHE[key]
