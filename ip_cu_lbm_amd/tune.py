"""Autotuner of the fused step's line axis.

The fused step gives one workgroup to each line of cells along its stream axis and one work-item
to each cell of the line, so the line axis fixes the workgroup's shape (the "tile") and the
number of workgroups.  On a grid with a stretched axis the lines must run along it; on a uniform
grid any axis with lines of 2-512 cells whose rows fit the LDS budget can carry them.  The
library's own choice is the axis whose lines fill their 64-wide wavefronts best (z on ties).  The
tuner instead times a few steps of every admissible axis on the device and keeps the fastest,
per grid shape and configuration (``shape_key``), in a cache that can be saved as JSON and
reused.  Cells per work-item stay one in the step: the line is the unit of the interpolated
transport and of the LDS exchange.

The reference has no counterpart (one CUDA thread per cell, fixed block size).
"""
from __future__ import annotations

import json
import os
from typing import Dict, Optional

from . import IplError, Simulation, _abi, cases

AXES = {1: "x", 2: "y", 3: "z"}


def shape_key(case: "cases.Case") -> str:
    """What the step's speed depends on besides the values: lattice, scalar field and lattice,
    storage, grid, boundaries and the optional terms."""
    body = case.body is not None and len(case.body.markers) > 0
    return (f"Q{case.lattice}-S{case.scalar}:{case.scalar_lattice}-st{case.storage}-n{'x'.join(map(str, case.n))}"
            f"-per{''.join(map(str, case.periodic))}-bc{case.wall_bc}-flux{case.flux_axes}-ib{int(body)}"
            f"-les{case.les}-wf{case.wall_function}")


def _time(case: "cases.Case", steps: int, warmup: int, library: Optional[str]) -> Optional[float]:
    try:
        sim = Simulation(case, library)
    except IplError:
        return None  # no lines along this axis (too long, too short, LDS) or not a fused case
    with sim:
        if sim.path != _abi.PATH_FUSED:
            return None
        sim.step(warmup)
        sim.prepare()
        return sim.step_timed(steps) / steps


def tune_line_axis(case: "cases.Case", steps: int = 100, warmup: int = 20, cache: Optional[Dict] = None,
                   library: Optional[str] = None) -> dict:
    """Times `steps` steps (after `warmup`) of `case` with its lines along each admissible axis
    and along the library's own choice; returns {"key", "line_axis" (1-3, the fastest),
    "ms_per_step": {axis: ms}, "auto_ms_per_step"} and stores it in `cache` under the shape key.
    The timed simulations are scratch copies: `case` itself is not stepped."""
    key = shape_key(case)
    if cache is not None and key in cache:
        return cache[key]
    times = {}
    for a in AXES:
        t = _time(case.replace(line_axis=a, path=_abi.PATH_FUSED), steps, warmup, library)
        if t is not None:
            times[a] = t
    if not times:
        raise ValueError(f"{case.name}: no line axis admits the fused path")
    auto = _time(case.replace(line_axis=0, path=_abi.PATH_FUSED), steps, warmup, library)
    best = min(times, key=times.get)
    res = {"key": key, "line_axis": best, "ms_per_step": {AXES[a]: t for a, t in times.items()},
           "auto_ms_per_step": auto}
    if cache is not None:
        cache[key] = res
    return res


def apply(case: "cases.Case", cache: Dict) -> "cases.Case":
    """`case` with the cached line axis of its shape (unchanged when the shape was not tuned)."""
    hit = cache.get(shape_key(case))
    return case.replace(line_axis=int(hit["line_axis"])) if hit else case


def load_cache(path: str) -> Dict:
    if not os.path.exists(path):
        return {}
    with open(path) as fh:
        return json.load(fh)


def save_cache(cache: Dict, path: str) -> None:
    with open(path, "w") as fh:
        json.dump(cache, fh, indent=1, sort_keys=True)
