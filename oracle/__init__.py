"""Python handle on the FP64 CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module,
and only to check or to time the reference algorithm on the CPU; the product never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_char_p, c_double, c_int, c_int32, c_int64, c_size_t, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_LIB = None
_MUTANTS = {}


def build_mutant(n: int) -> str:
    """The oracle built with planted mistake number n (ORC_MUTANT, see oracle.cpp): test
    infrastructure for showing that a pin fails on that mistake."""
    out = os.path.join(_HERE, "_build", f"liboracle_mutant{n}.so")
    src = os.path.join(_HERE, "oracle.cpp")
    if not os.path.exists(out) or os.path.getmtime(out) < os.path.getmtime(src):
        os.makedirs(os.path.dirname(out), exist_ok=True)
        subprocess.run(["g++", "-O2", "-fopenmp", "-fPIC", "-std=c++17", "-shared", f"-DORC_MUTANT={int(n)}", src,
                        "-o", out + ".tmp"], check=True)
        os.replace(out + ".tmp", out)
    return out


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++ -fopenmp)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return LIB_PATH


def load(path: str = None):
    """The oracle library (the normal build, or `path`, e.g. a mutant of build_mutant)."""
    global _LIB
    if path is not None and path in _MUTANTS:
        return _MUTANTS[path]
    if path is not None or _LIB is None:
        if path is None and not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(path or LIB_PATH)
        DP = POINTER(c_double)
        lib.orc_create.argtypes = [c_void_p, c_void_p, c_char_p, c_int]
        lib.orc_create.restype = c_void_p
        lib.orc_destroy.argtypes = [c_void_p]
        lib.orc_step.argtypes = [c_void_p, c_int]
        lib.orc_time_steps.argtypes = [c_void_p, c_int]
        lib.orc_time_steps.restype = c_double
        lib.orc_stage.argtypes = [c_void_p, c_int]
        lib.orc_ib_stage.argtypes = [c_void_p, c_int]
        lib.orc_ib_stage.restype = c_int
        lib.orc_info.argtypes = [c_void_p, POINTER(c_int64), POINTER(c_int64)]
        lib.orc_read.argtypes = [c_void_p, c_int, DP, c_size_t]
        lib.orc_write.argtypes = [c_void_p, c_int, DP, c_size_t]
        lib.orc_global_sums.argtypes = [c_void_p, DP]
        lib.orc_table.argtypes = [c_void_p, c_int, POINTER(c_int32), c_int64]
        lib.orc_table.restype = c_int64
        lib.orc_lagrange.argtypes = [DP, c_double, c_int, DP]
        lib.orc_mrt.argtypes = [c_int, DP, DP, POINTER(c_int32), POINTER(c_int32)]
        lib.orc_lattice.argtypes = [c_int, POINTER(c_int32), DP]
        lib.orc_ib_attach.argtypes = [c_void_p, c_void_p, c_char_p, c_int]
        lib.orc_ib_body_state.argtypes = [c_void_p, DP]
        lib.orc_ib_read_markers.argtypes = [c_void_p, c_int, DP, c_size_t]
        lib.orc_ib_peskin4.argtypes = [c_double]
        lib.orc_ib_peskin4.restype = c_double
        lib.orc_ib_rotation.argtypes = [DP, DP]
        lib.orc_set_open_boundary.argtypes = [c_void_p, POINTER(c_int32), c_int, DP, DP, DP, DP]
        lib.orc_hydrostatic_pressure.argtypes = [c_void_p, c_int, c_double, DP]
        lib.orc_wall_cells.argtypes = [c_void_p, c_int, POINTER(c_int32)]
        lib.orc_wall_cells.restype = c_int
        lib.orc_change_dt.argtypes = [c_void_p, c_double]
        lib.orc_les.argtypes = [c_void_p]
        lib.orc_spalding_utau.argtypes = [c_double, c_double, c_double]
        lib.orc_spalding_utau.restype = c_double
        if path is not None:
            _MUTANTS[path] = lib
            return lib
        _LIB = lib
    return _LIB


def _dp(a):
    return a.ctypes.data_as(POINTER(c_double))


class Oracle:
    """FP64 reference-algorithm run of a Case (same inputs as the HIP library)."""

    def __init__(self, case, lib_path: str = None):
        lib = load(lib_path)
        cfg, ini, keep = case.to_c()
        err = ctypes.create_string_buffer(256)
        h = lib.orc_create(ctypes.byref(cfg), ctypes.byref(ini), err, 256)
        if not h:
            raise ValueError(f"oracle rejected the case: {err.value.decode()}")
        self._h = h
        self._lib = lib
        self.case = case
        n, g = c_int64(), c_int64()
        lib.orc_info(h, ctypes.byref(n), ctypes.byref(g))
        self.n_calc, self.n_ghost = n.value, g.value
        del keep
        if getattr(case, "body", None) is not None:
            self.ib_attach(case.body)

    def ib_attach(self, body):
        cfg, keep = body.to_c()
        err = ctypes.create_string_buffer(256)
        rc = self._lib.orc_ib_attach(self._h, ctypes.byref(cfg), err, 256)
        if rc != 0:
            raise ValueError(f"oracle rejected the body: {err.value.decode()}")
        self._n_markers = len(body.markers)
        del keep

    def body_state(self) -> dict:
        out = np.zeros(19)
        self._lib.orc_ib_body_state(self._h, _dp(out))
        return {"pos": out[0:3], "vel": out[3:6], "omega": out[6:9], "quat": out[9:13], "force": out[13:16],
                "torque": out[16:19]}

    def markers(self, which: str) -> np.ndarray:
        from ip_cu_lbm_amd._abi import IB_ARRAYS
        m = self._n_markers
        out = np.empty(m if which == "node" else 3 * m)
        rc = self._lib.orc_ib_read_markers(self._h, IB_ARRAYS[which], _dp(out), out.size)
        if rc != 0:
            raise ValueError(f"oracle marker read failed ({rc})")
        return out if which == "node" else out.reshape(m, 3)

    def close(self):
        if self._h:
            self._lib.orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n: int = 1):
        self._lib.orc_step(self._h, n)

    def time_steps(self, n: int) -> float:
        return self._lib.orc_time_steps(self._h, n)

    def stage(self, name_or_id):
        from ip_cu_lbm_amd._abi import STAGES
        sid = STAGES[name_or_id] if isinstance(name_or_id, str) else int(name_or_id)
        self._lib.orc_stage(self._h, sid)

    def ib_stage(self, name: str):
        from ip_cu_lbm_amd._abi import IB_STAGES
        rc = self._lib.orc_ib_stage(self._h, IB_STAGES[name])
        if rc != 0:
            raise ValueError(f"oracle IB stage {name} failed ({rc})")

    def read(self, field: str) -> np.ndarray:
        from ip_cu_lbm_amd._abi import FIELDS
        fid = FIELDS[field]
        if field in ("f", "g"):
            q = self.case.lattice if field == "f" else (self.case.scalar_lattice or self.case.lattice)
            out = np.empty(q * self.n_calc)
        else:
            out = np.empty(self.n_calc)
        rc = self._lib.orc_read(self._h, fid, _dp(out), out.size)
        if rc != 0:
            raise ValueError(f"oracle read of {field} failed ({rc})")
        return out.reshape(-1, self.n_calc) if field in ("f", "g") else out

    def write(self, field: str, data: np.ndarray):
        from ip_cu_lbm_amd._abi import FIELDS
        a = np.ascontiguousarray(data, dtype=np.float64).ravel()
        rc = self._lib.orc_write(self._h, FIELDS[field], _dp(a), a.size)
        if rc != 0:
            raise ValueError(f"oracle write of {field} failed ({rc})")

    def global_sums(self) -> np.ndarray:
        out = np.zeros(3)
        self._lib.orc_global_sums(self._h, _dp(out))
        return out

    def set_open_boundary(self, cells, p, u, v, w):
        cells = np.ascontiguousarray(cells, dtype=np.int32).ravel()
        n = cells.size
        vals = [np.ascontiguousarray(np.broadcast_to(np.asarray(x, dtype=np.float64), (n,))) for x in (p, u, v, w)]
        rc = self._lib.orc_set_open_boundary(self._h, cells.ctypes.data_as(POINTER(c_int32)), n, *[_dp(x) for x in vals])
        if rc != 0:
            raise ValueError("open-boundary cell out of range")

    def hydrostatic_pressure(self, boussinesq: bool, center_z: float = 0.0) -> np.ndarray:
        out = np.empty(self.n_calc)
        self._lib.orc_hydrostatic_pressure(self._h, 1 if boussinesq else 0, float(center_z), _dp(out))
        return out

    def les(self):
        self._lib.orc_les(self._h)

    def change_dt(self, alpha: float):
        self._lib.orc_change_dt(self._h, float(alpha))

    def wall_cells(self, face: int) -> np.ndarray:
        out = np.empty(max(self.n_calc, 1), dtype=np.int32)
        m = self._lib.orc_wall_cells(self._h, int(face), out.ctypes.data_as(POINTER(c_int32)))
        return out[:m]

    def table(self, which: int) -> np.ndarray:
        n = self._lib.orc_table(self._h, which, None, 0)
        out = np.empty(n, dtype=np.int32)
        self._lib.orc_table(self._h, which, out.ctypes.data_as(POINTER(c_int32)), n)
        return out


def spalding_utau(abs_u: float, y_dist: float, nu: float) -> float:
    return load().orc_spalding_utau(float(abs_u), float(y_dist), float(nu))


def peskin4(r: float) -> float:
    return load().orc_ib_peskin4(float(r))


def rotation(q) -> np.ndarray:
    q = np.ascontiguousarray(q, dtype=np.float64)
    R = np.empty(9)
    load().orc_ib_rotation(_dp(q), _dp(R))
    return R.reshape(3, 3)


def lagrange(x, xa):
    lib = load()
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty(len(x))
    lib.orc_lagrange(_dp(x), float(xa), len(x), _dp(out))
    return out


def mrt(Q: int):
    lib = load()
    M = np.empty(Q * Q)
    Mi = np.empty(Q * Q)
    st = np.empty(8, dtype=np.int32)
    ns = c_int32()
    lib.orc_mrt(Q, _dp(M), _dp(Mi), st.ctypes.data_as(POINTER(c_int32)), ctypes.byref(ns))
    return M.reshape(Q, Q), Mi.reshape(Q, Q), st[: ns.value].tolist()


def lattice(Q: int):
    lib = load()
    c = np.empty(3 * Q, dtype=np.int32)
    w = np.empty(Q)
    lib.orc_lattice(Q, c.ctypes.data_as(POINTER(c_int32)), _dp(w))
    return c.reshape(Q, 3), w
