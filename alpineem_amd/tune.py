"""Choice of the bulk launches' z chunk ("planes") by measurement.

A bulk thread marches a column of z planes through its work-group's tile.
How many planes it takes sets the number of work-groups and the length of
each thread's loop. The best value depends on the grid, and for a batch on
how many simulations share each launch. The value changes no result: every
cell runs the same FP32 operations in the same order whatever the chunk is,
so it can be chosen by timing throwaway runs. The chosen value is then
passed as ``planes=`` to the Simulations that do the work.

The reference has no counterpart: its update loops are fixed
OpenMP/OpenACC nests over the whole grid (main_fdtd/fdtd_solver.f90:5670-5702).
"""
from typing import Dict, Sequence, Tuple

from .solver import Batch, Plan, Simulation

# 0 is the library's own choice: about 10 planes per thread in balanced
# chunks, fewer on small grids and thin boundary boxes.
CANDIDATES = (0, 2, 4, 6, 8, 10, 16)


def _check(plans: Sequence[Plan], candidates: Sequence[int], steps: int, warmup: int) -> None:
    if not plans:
        raise ValueError("no plans to tune on")
    if not candidates or any(int(c) < 0 for c in candidates):
        raise ValueError(f"candidates must be a non-empty list of planes >= 0, got {list(candidates)}")
    if steps < 1 or warmup < 1:
        raise ValueError("steps and warmup must be at least 1 (warmup absorbs the graph capture)")
    short = min(p.time_steps for p in plans)
    if warmup + steps > short:
        raise ValueError(f"warmup + steps = {warmup + steps} exceeds the shortest deck's time_steps ({short})")


def tune_planes(plan: Plan, candidates: Sequence[int] = CANDIDATES, steps: int = 20, warmup: int = 5,
                **options) -> Tuple[int, Dict[int, float]]:
    """For each candidate, time ``steps`` graph-launched steps of a throwaway
    Simulation of ``plan`` after ``warmup`` steps. ``options`` are passed to
    Simulation. Returns (fastest candidate, {candidate: ms per step})."""
    _check([plan], candidates, steps, warmup)
    ms: Dict[int, float] = {}
    for c in candidates:
        sim = Simulation(plan, planes=int(c), **options)
        try:
            sim.run(warmup)
            ms[int(c)] = sim.run(steps)
        finally:
            sim.close()
    return min(ms, key=ms.get), ms


def tune_batch_planes(plans: Sequence[Plan], candidates: Sequence[int] = CANDIDATES, steps: int = 20,
                      warmup: int = 5, **options) -> Tuple[int, Dict[int, float]]:
    """As tune_planes, for one Batch of throwaway Simulations (one per plan,
    all with the same candidate). Returns (fastest candidate, {candidate: ms
    per batched step})."""
    _check(plans, candidates, steps, warmup)
    ms: Dict[int, float] = {}
    for c in candidates:
        sims = [Simulation(p, planes=int(c), **options) for p in plans]
        batch = Batch(sims)
        try:
            batch.run(warmup)
            ms[int(c)] = batch.run(steps)
        finally:
            batch.close()
            for s in sims:
                s.close()
    return min(ms, key=ms.get), ms
