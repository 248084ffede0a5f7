"""MI355X-native interpolation-based lattice Boltzmann step (HIP, gfx950).

Host-side mirror of the reference's run loop (init -> N x step -> readout, e.g.
LockExchange/LockExchange.cu:12-291) over the C ABI of include/ipl.h:

    from ip_cu_lbm_amd import Simulation, cases
    sim = Simulation(cases.lock_exchange(storage=F32))
    sim.step(100)
    u = sim.read("u")

Every compute call goes to the HIP library ip_cu_lbm_amd/_lib/libipl.so; there is no CPU
fallback, and the constructor raises if the library is missing.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _abi, cases
from ._abi import (BC_BOUNCE_BACK, BC_SLIP, COLL_MRT, COLL_SRT, F16, F32, F64, FIELDS, FLUX_NONE, IB_DIRECT,
                   IB_KERNEL_PESKIN4, IB_KERNEL_REFERENCE, IB_STRESS,
                   FLUX_REFERENCE, FLUX_X, FLUX_Y, FLUX_Z, PATH_AUTO, PATH_FUSED, PATH_PLAIN,
                   SCALAR_NONE, SCALAR_PHASE_FIELD, SCALAR_SALINITY, STAGES, IplError)

__all__ = ["Simulation", "Batch", "estimate_device_bytes", "hydrostatic_pressure", "wall_cells", "cases", "IplError", "F64", "F32", "F16", "PATH_AUTO",
           "PATH_PLAIN", "PATH_FUSED"]


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Simulation:
    """One simulation context on the current GPU (ipl_create ... ipl_destroy)."""

    def __init__(self, case: "cases.Case", library: str = None):
        self._lib = _abi.load_library(library)
        self.case = case
        cfg, ini, keep = case.to_c()
        h = ctypes.c_void_p()
        rc = self._lib.ipl_create(ctypes.byref(cfg), ctypes.byref(ini), ctypes.byref(h))
        del keep
        if rc != 0:
            msg = self._lib.ipl_last_error_message(None).decode()
            raise IplError(rc, f"{self._lib.ipl_error_string(rc).decode()}: {msg}")
        self._h = h
        n, g, b, p = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32()
        self._lib.ipl_info(h, ctypes.byref(n), ctypes.byref(g), ctypes.byref(b), ctypes.byref(p))
        self.n_calc, self.n_ghost, self.device_bytes, self.path = n.value, g.value, b.value, p.value
        # why the plain path runs this case ("" on the fused path)
        self.path_reason = self._lib.ipl_path_reason(h).decode()
        if case.body is not None:
            self.ib_attach(case.body)
            self._lib.ipl_info(h, None, None, ctypes.byref(b), None)
            self.device_bytes = b.value

    # ---- lifecycle --------------------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._lib.ipl_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            msg = self._lib.ipl_last_error_message(self._h).decode()
            raise IplError(rc, f"{self._lib.ipl_error_string(rc).decode()}: {msg}")

    # ---- run loop ---------------------------------------------------------------------------
    def step(self, n: int = 1):
        self._check(self._lib.ipl_step(self._h, int(n)))

    def set_graph(self, on: bool):
        """Replay captured HIP graphs of the step (default on) or launch every step directly."""
        self._check(self._lib.ipl_set_graph(self._h, 1 if on else 0))

    def prepare(self):
        """Captures the step graphs now (recorded, not run), so later steps pay no capture."""
        self._check(self._lib.ipl_prepare(self._h))

    def set_shared(self, on: bool):
        """The context shares the device with other contexts stepping at the same time: an
        immersed body's step tail stays on the context's stream (set before prepare())."""
        self._check(self._lib.ipl_set_shared(self._h, 1 if on else 0))

    def graph_captures(self) -> int:
        """Number of graph captures this context has made."""
        n = ctypes.c_int32()
        self._check(self._lib.ipl_graph_captures(self._h, ctypes.byref(n)))
        return n.value

    def set_open_boundary(self, cells, p, u, v, w):
        """Inflow / outflow cells (canonical indices) reset to feq(p, u, v, w) after every step
        (set_velocity1 + set_wall_boundary1); scalars or per-cell arrays; empty cells remove it."""
        cells = np.ascontiguousarray(cells, dtype=np.int32).ravel()
        n = cells.size
        vals = [np.ascontiguousarray(np.broadcast_to(np.asarray(x, dtype=np.float64), (n,))) for x in (p, u, v, w)]
        self._check(self._lib.ipl_set_open_boundary(self._h, cells.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), n,
                                                    *[_dp(x) for x in vals]))

    def change_dt(self, alpha: float):
        """dt / alpha, c * alpha, p / alpha^2, f reset to feq(p, u) (reset_items_dt + reset_pressure)."""
        self._check(self._lib.ipl_change_dt(self._h, float(alpha)))

    def write_fields_csv(self, path: str):
        """The reference's output() file: num_calc line, 15-column header, one row per cell."""
        self._check(self._lib.ipl_write_fields_csv(self._h, os.fsencode(path)))

    def write_markers_csv(self, path: str):
        """The reference's IB_csv file: marker positions, velocities and forces."""
        self._check(self._lib.ipl_write_markers_csv(self._h, os.fsencode(path)))

    def write_vtu(self, path: str):
        """VTU of hexahedral cells with the cell data of make_VTK.py."""
        self._check(self._lib.ipl_write_vtu(self._h, os.fsencode(path)))

    def step_timed(self, n: int) -> float:
        """Runs n steps and returns their device time in milliseconds (HIP events)."""
        ms = ctypes.c_double()
        self._check(self._lib.ipl_step_timed(self._h, int(n), ctypes.byref(ms)))
        return ms.value

    def profile_steps(self, n: int):
        """Runs n steps with per-launch HIP-event timing; returns (dominant kernel ms per launch,
        other kernels ms per step, total ms)."""
        out = np.zeros(3)
        self._check(self._lib.ipl_profile_steps(self._h, int(n), _dp(out)))
        return float(out[0]), float(out[1]), float(out[2])

    def stage(self, name):
        sid = STAGES[name] if isinstance(name, str) else int(name)
        self._check(self._lib.ipl_stage(self._h, sid))

    # ---- readout ----------------------------------------------------------------------------
    def _qf(self, field):
        if field == "f":
            return self.case.lattice
        return self.case.scalar_lattice or self.case.lattice

    def read(self, field: str) -> np.ndarray:
        if field in ("f", "g"):
            out = np.empty(self._qf(field) * self.n_calc)
        else:
            out = np.empty(self.n_calc)
        self._check(self._lib.ipl_read(self._h, FIELDS[field], _dp(out), out.size))
        return out.reshape(-1, self.n_calc) if field in ("f", "g") else out

    def write(self, field: str, data):
        a = np.ascontiguousarray(data, dtype=np.float64).ravel()
        self._check(self._lib.ipl_write(self._h, FIELDS[field], _dp(a), a.size))

    def global_sums(self) -> np.ndarray:
        out = np.zeros(3)
        self._check(self._lib.ipl_global_sums(self._h, _dp(out)))
        return out

    def global_sums_noalloc(self) -> np.ndarray:
        """The same sums without a device allocation (the library's per-GPU scratch array)."""
        out = np.zeros(3)
        self._check(self._lib.ipl_global_sums_noalloc(self._h, _dp(out)))
        return out

    # ---- immersed boundary ------------------------------------------------------------------
    def ib_attach(self, body: "cases.Body"):
        cfg, keep = body.to_c()
        self._check(self._lib.ipl_ib_attach(self._h, ctypes.byref(cfg)))
        self._n_markers = len(body.markers)
        del keep

    def ib_stage(self, name: str):
        """One stage of the immersed-boundary step: 'flow' (a step without its IB sub-step), then
        'spm', 'force', 'update_vel', 'body', 'points' in the drivers' order (ipl_ib_stage)."""
        self._check(self._lib.ipl_ib_stage(self._h, _abi.IB_STAGES[name]))

    def body_state(self) -> dict:
        """FP64 body state: pos, vel, omega (body frame), quat, force, torque."""
        out = np.zeros(19)
        self._check(self._lib.ipl_ib_body_state(self._h, _dp(out)))
        return {"pos": out[0:3], "vel": out[3:6], "omega": out[6:9], "quat": out[9:13], "force": out[13:16],
                "torque": out[16:19]}

    def markers(self, which: str) -> np.ndarray:
        """Marker arrays (IB_csv): 'pos', 'vel', 'force', 'normal' (M x 3) or 'node' (M)."""
        m = self._n_markers
        out = np.empty(m if which == "node" else 3 * m)
        self._check(self._lib.ipl_ib_read_markers(self._h, _abi.IB_ARRAYS[which], _dp(out), out.size))
        return out if which == "node" else out.reshape(m, 3)


class Batch:
    """Independent simulations advanced together (ipl_batch_*).  batched=True (default): members
    whose step is the fused kernel alone advance in batched launches, one per kernel variant and
    step; the others (and every member with batched=False) launch their steps on their own
    streams before the batch waits for all of them."""

    def __init__(self, library: str = None, batched: bool = True, max_lines: int = None):
        self._lib = _abi.load_library(library)
        h = ctypes.c_void_p()
        rc = self._lib.ipl_batch_create(ctypes.byref(h))
        if rc != 0:
            raise IplError(rc, "ipl_batch_create")
        self._h = h
        self.members = []
        rc = self._lib.ipl_batch_set_launch(self._h, 1 if batched else 0)
        if rc != 0:
            raise IplError(rc, "ipl_batch_set_launch")
        if max_lines is not None:  # members with more lines step on their own streams
            rc = self._lib.ipl_batch_set_max_lines(self._h, int(max_lines))
            if rc != 0:
                raise IplError(rc, "ipl_batch_set_max_lines")

    def counters(self) -> dict:
        """Totals since creation: batched kernel launches, member-steps they advanced, member-steps
        launched on the members' own streams, graphs captured for the batched launches."""
        out = (ctypes.c_int64 * 4)()
        rc = self._lib.ipl_batch_counters(self._h, out)
        if rc != 0:
            raise IplError(rc, "ipl_batch_counters")
        return {"launches": out[0], "batched_steps": out[1], "own_steps": out[2], "graph_captures": out[3]}

    def graph_captures(self) -> int:
        return self.counters()["graph_captures"]

    def add(self, sim: Simulation):
        rc = self._lib.ipl_batch_add(self._h, sim._h)
        if rc != 0:
            raise IplError(rc, "ipl_batch_add")
        self.members.append(sim)

    def remove(self, sim: Simulation):
        rc = self._lib.ipl_batch_remove(self._h, sim._h)
        if rc != 0:
            raise IplError(rc, "ipl_batch_remove")
        self.members.remove(sim)

    def step(self, counts):
        """Advances member i by counts[i] steps (an int: every member by that many)."""
        if isinstance(counts, int):
            arr = (ctypes.c_int32 * 1)(counts)
            rc = self._lib.ipl_batch_step(self._h, arr, 1)
        else:
            arr = (ctypes.c_int32 * len(counts))(*[int(c) for c in counts])
            rc = self._lib.ipl_batch_step(self._h, arr, 0)
        if rc != 0:
            msg = self._lib.ipl_batch_last_error_message(self._h).decode()
            raise IplError(rc, f"{self._lib.ipl_error_string(rc).decode()}: {msg}")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.ipl_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def hydrostatic_pressure(case: "cases.Case", boussinesq: bool, center_z: float = 0.0, library: str = None) -> np.ndarray:
    """The initial hydrostatic p of every calc cell (hydrostatic_pressure, basic.cu:239-252)."""
    lib = _abi.load_library(library)
    cfg, _ini, keep = case.to_c()
    out = np.empty(case.n_cells)
    rc = lib.ipl_hydrostatic_pressure(ctypes.byref(cfg), 1 if boussinesq else 0, float(center_z), _dp(out), out.size)
    del keep
    if rc != 0:
        raise IplError(rc, "ipl_hydrostatic_pressure")
    return out


def wall_cells(case: "cases.Case", face: int, library: str = None) -> np.ndarray:
    """Calc cells next to a wall face (set_walln, set_bound.cu:8-23): 1 +x, 2 +y, 3 -x, 4 -y, 5 +z, 6 -z."""
    lib = _abi.load_library(library)
    cfg, _ini, keep = case.to_c()
    n = ctypes.c_int64()
    rc = lib.ipl_wall_cells(ctypes.byref(cfg), int(face), None, ctypes.byref(n))
    out = np.empty(max(n.value, 1), dtype=np.int32)
    if rc == 0:
        rc = lib.ipl_wall_cells(ctypes.byref(cfg), int(face), out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                ctypes.byref(n))
    del keep
    if rc != 0:
        raise IplError(rc, "ipl_wall_cells")
    return out[:n.value]


def estimate_device_bytes(case: "cases.Case", library: str = None) -> int:
    """Device bytes ipl_create would allocate for the case (admission against a budget)."""
    lib = _abi.load_library(library)
    cfg, _ini, keep = case.to_c()
    out = ctypes.c_int64()
    rc = lib.ipl_estimate_device_bytes(ctypes.byref(cfg), ctypes.byref(out))
    total = int(out.value)
    if rc == 0 and getattr(case, "body", None) is not None:
        bcfg, bkeep = case.body.to_c()
        rc = lib.ipl_estimate_ib_bytes(ctypes.byref(cfg), ctypes.byref(bcfg), ctypes.byref(out))
        total += int(out.value)
        del bkeep
    del keep
    if rc != 0:
        raise IplError(rc, "ipl_estimate_device_bytes")
    return total


def write_cd_csv(path: str, c_d, time, library: str = None):
    """The drivers' C_D.csv (out_C_D, SphereFall3D/SphereFall.cu:32-42): header "C_D,C_time" and one
    row per record, e.g. the settling velocity or a drag coefficient at each save interval."""
    lib = _abi.load_library(library)
    a = np.ascontiguousarray(c_d, dtype=np.float64).ravel()
    t = np.ascontiguousarray(time, dtype=np.float64).ravel()
    if a.size != t.size:
        raise ValueError("C_D and time records differ in length")
    rc = lib.ipl_write_cd_csv(os.fsencode(path), _dp(a), _dp(t), a.size)
    if rc != 0:
        raise IplError(rc, f"ipl_write_cd_csv {path}")


def settling_history(sim: "Simulation", steps: int, every: int, dt: float):
    """Runs `steps` steps and records, every `every` steps, the body's z velocity and the time, as
    the falling-sphere driver does for its C_D.csv (SphereFall3D/SphereFall.cu:424, 437)."""
    vel, times = [], []
    done = 0
    while done < steps:
        k = min(every, steps - done)
        sim.step(k)
        done += k
        vel.append(float(sim.body_state()["vel"][2]))
        times.append(done * dt)
    return np.array(vel), np.array(times)
