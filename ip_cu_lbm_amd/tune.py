"""Autotuner of the fused step's tile: its line axis and its workgroup size.

The fused step gives one work-item to each cell of a line of cells along its stream axis, and a
workgroup holds one or more consecutive lines, each in a slot of whole 64-wide wavefronts
(``workgroup`` work-items, as many slots as fit), so the line axis and the workgroup size fix the
tile and the number of workgroups.  On a grid with a stretched axis the lines must run along it;
on a uniform grid any axis with lines of 2-512 cells whose rows fit the LDS budget can carry them.
The library's own choice is one line per workgroup, along the axis whose lines fill their
wavefronts best (z on ties).  The tuner instead times a few steps of every admissible (axis, workgroup) pair
on the device and keeps the fastest, per grid shape and configuration (``shape_key``), in a cache
that can be saved as JSON and reused.  Cells per work-item stay one in the step: the line is the
unit of the interpolated transport and of the LDS exchange.  Every tile gives the same results
bit for bit along one axis (the lines are independent; only their grouping changes).

The reference has no counterpart (one CUDA thread per cell, blocks of 64 threads).
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


WORKGROUPS = (64, 128, 192, 256, 320, 384, 448, 512)


def workgroups_for(n: int) -> list:
    """The workgroup sizes the tuner tries for lines of n cells: 0 (the library's choice, one line)
    and, for every further count of line slots (n rounded up to 64 work-items) per workgroup, the
    smallest size that holds it."""
    slot = -(-n // 64) * 64
    out, seen = [0], {1}
    for w in WORKGROUPS:
        if w >= slot and w // slot not in seen:
            seen.add(w // slot)
            out.append(w)
    return out


def tune_line_axis(case: "cases.Case", steps: int = 100, warmup: int = 20, cache: Optional[Dict] = None,
                   library: Optional[str] = None, workgroups: bool = True) -> dict:
    """Times `steps` steps (after `warmup`) of `case` with its lines along each admissible axis,
    each with the workgroup sizes of ``workgroups_for`` (only the library's choice when
    `workgroups` is False), and the library's own tile; returns {"key", "line_axis" (1-3) and
    "workgroup" (0 or work-items) of the fastest, "ms_per_step": {"axis" or "axis/workgroup": ms},
    "auto_ms_per_step"} and stores it in `cache` under the shape key.  The timed simulations are
    scratch copies: `case` itself is not stepped."""
    key = shape_key(case)
    if cache is not None and key in cache:
        return cache[key]
    times = {}
    for a in AXES:
        for w in (workgroups_for(case.n[a - 1]) if workgroups else [0]):
            t = _time(case.replace(line_axis=a, workgroup=w, path=_abi.PATH_FUSED), steps, warmup, library)
            if t is not None:
                times[(a, w)] = t
    if not times:
        raise ValueError(f"{case.name}: no line axis admits the fused path")
    auto = _time(case.replace(line_axis=0, workgroup=0, path=_abi.PATH_FUSED), steps, warmup, library)
    best = min(times, key=times.get)
    res = {"key": key, "line_axis": best[0], "workgroup": best[1],
           "ms_per_step": {(AXES[a] if w == 0 else f"{AXES[a]}/{w}"): t for (a, w), t in times.items()},
           "auto_ms_per_step": auto}
    if cache is not None:
        cache[key] = res
    return res


def apply(case: "cases.Case", cache: Dict) -> "cases.Case":
    """`case` with the cached tile of its shape (unchanged when the shape was not tuned)."""
    hit = cache.get(shape_key(case))
    if not hit:
        return case
    return case.replace(line_axis=int(hit["line_axis"]), workgroup=int(hit.get("workgroup", 0)))


def load_cache(path: str) -> Dict:
    if not os.path.exists(path):
        return {}
    with open(path) as fh:
        return json.load(fh)


def save_cache(cache: Dict, path: str) -> None:
    with open(path, "w") as fh:
        json.dump(cache, fh, indent=1, sort_keys=True)
