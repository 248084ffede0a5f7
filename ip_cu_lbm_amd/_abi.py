"""ctypes mirror of include/ipl.h and the loader of the HIP library.

The layout of ``IplConfig`` / ``IplInit`` must match ``ipl_config`` / ``ipl_init`` field by field;
tests/test_abi.py checks the sizes against the C compiler's.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_int32, c_int64, c_size_t, c_uint8, c_void_p, c_char_p

IPL_OK = 0
IPL_ERR_INVALID_CONFIG = 1
IPL_ERR_HIP = 2
IPL_ERR_OUT_OF_MEMORY = 3
IPL_ERR_NONFINITE = 4
IPL_ERR_UNSUPPORTED = 5
IPL_ERR_ARGUMENT = 6
IPL_ERR_BUDGET = 7

BC_SLIP, BC_BOUNCE_BACK = 0, 1
COLL_MRT, COLL_SRT = 0, 1
SCALAR_NONE, SCALAR_SALINITY, SCALAR_PHASE_FIELD = 0, 1, 2
F64, F32, F16 = 0, 1, 2
PATH_AUTO, PATH_PLAIN, PATH_FUSED = 0, 1, 2
FLUX_NONE, FLUX_X, FLUX_Y, FLUX_Z, FLUX_REFERENCE = 0, 1, 2, 4, -1

FIELDS = {
    "p": 0, "u": 1, "v": 2, "w": 3, "rho": 4, "S": 5, "phi": 6,
    "Fx": 7, "Fy": 8, "Fz": 9, "tau": 10, "taus": 11, "f": 12, "g": 13,
}

IB_STAGES = {"flow": 0, "spm": 1, "force": 2, "update_vel": 3, "body": 4, "points": 5}

STAGES = {
    "equ_f": 0, "Force": 1, "collide": 2, "IP_f": 3, "scalar_collide": 4, "IP_g": 5,
    "update_scalar": 6, "update_rho": 7, "set_wall_rho": 8, "resetF": 9, "equ_g": 10,
}

_DP = POINTER(c_double)


class IplConfig(ctypes.Structure):
    _fields_ = [
        ("lattice", c_int32), ("nx", c_int32), ("ny", c_int32), ("nz", c_int32),
        ("dz", c_double), ("dt", c_double),
        ("pos", _DP * 3), ("width", _DP * 3),
        ("solid", POINTER(c_uint8)),
        ("periodic", c_int32 * 3), ("wall_bc", c_int32), ("flux_axes", c_int32),
        ("stencil", c_int32), ("collision", c_int32),
        ("s_rates", c_double * 27), ("tau", c_double), ("taus", c_double),
        ("rho_L", c_double), ("rho_H", c_double), ("beta_s", c_double), ("gravity", c_double),
        ("boussinesq", c_int32), ("sigma", c_double), ("pf_thickness", c_double),
        ("scalar", c_int32), ("scalar_lattice", c_int32), ("storage", c_int32), ("path", c_int32),
        ("f16_scale", c_double), ("check_finite", c_int32), ("pf_tau", c_double),
        ("wall_function", c_int32), ("workgroup", c_int32), ("wall_nu", c_double),
        ("les", c_int32), ("line_axis", c_int32), ("mu_L", c_double), ("mu_H", c_double),
    ]


class IplInit(ctypes.Structure):
    _fields_ = [(n, _DP) for n in ("p", "u", "v", "w", "S", "phi", "tau", "taus")]


IB_STRESS, IB_DIRECT = 0, 1
IB_KERNEL_REFERENCE, IB_KERNEL_PESKIN4 = 0, 1
IB_ARRAYS = {"pos": 0, "vel": 1, "force": 2, "normal": 3, "node": 4}


class IplIbConfig(ctypes.Structure):
    _fields_ = [
        ("method", c_int32), ("kernel", c_int32), ("iterations", c_int32), ("moving", c_int32),
        ("n_markers", c_int32), ("marker_pos", _DP), ("marker_normal", _DP),
        ("spacing", c_double), ("rho_f", c_double), ("density", c_double), ("mass", c_double),
        ("inertia", c_double * 3), ("gravity", c_double * 3),
        ("pos", c_double * 3), ("vel", c_double * 3), ("omega", c_double * 3), ("quat", c_double * 4),
        ("spm", c_int32), ("reserved0", c_int32), ("spm_a", c_double), ("spm_b", c_double), ("reserved", c_int32 * 2),
    ]


def lib_path() -> str:
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libipl.so")


_LIBS = {}


def load_library(path: str = None):
    """Load the HIP library.  Raises if it has not been built: there is no CPU fallback.
    `path` lets the test suite load the CPU emulation build of the same sources."""
    path = path or lib_path()
    if path in _LIBS:
        return _LIBS[path]
    if not os.path.exists(path):
        raise RuntimeError(
            f"HIP library {path} is not built; run `python -m ip_cu_lbm_amd.build` "
            "(hipcc --offload-arch=gfx950)")
    lib = ctypes.CDLL(path)
    vp = c_void_p
    lib.ipl_create.argtypes = [POINTER(IplConfig), POINTER(IplInit), POINTER(vp)]
    lib.ipl_create.restype = c_int32
    lib.ipl_step.argtypes = [vp, c_int32]
    lib.ipl_step.restype = c_int32
    lib.ipl_step_timed.argtypes = [vp, c_int32, POINTER(c_double)]
    lib.ipl_step_timed.restype = c_int32
    lib.ipl_read.argtypes = [vp, c_int32, _DP, c_size_t]
    lib.ipl_read.restype = c_int32
    lib.ipl_write.argtypes = [vp, c_int32, _DP, c_size_t]
    lib.ipl_write.restype = c_int32
    lib.ipl_global_sums.argtypes = [vp, _DP]
    lib.ipl_global_sums.restype = c_int32
    lib.ipl_global_sums_noalloc.argtypes = [vp, _DP]
    lib.ipl_global_sums_noalloc.restype = c_int32
    lib.ipl_info.argtypes = [vp, POINTER(c_int64), POINTER(c_int64), POINTER(c_int64), POINTER(c_int32)]
    lib.ipl_info.restype = c_int32
    lib.ipl_destroy.argtypes = [vp]
    lib.ipl_destroy.restype = c_int32
    lib.ipl_error_string.argtypes = [c_int32]
    lib.ipl_error_string.restype = c_char_p
    lib.ipl_last_error_message.argtypes = [vp]
    lib.ipl_last_error_message.restype = c_char_p
    lib.ipl_path_reason.argtypes = [vp]
    lib.ipl_path_reason.restype = c_char_p
    lib.ipl_profile_steps.argtypes = [vp, c_int32, _DP]
    lib.ipl_profile_steps.restype = c_int32
    lib.ipl_stage.argtypes = [vp, c_int32]
    lib.ipl_stage.restype = c_int32
    lib.ipl_build_tables.argtypes = [POINTER(IplConfig), c_int32, c_void_p, POINTER(c_int64)]
    lib.ipl_build_tables.restype = c_int32
    lib.ipl_ib_attach.argtypes = [vp, POINTER(IplIbConfig)]
    lib.ipl_ib_attach.restype = c_int32
    lib.ipl_ib_stage.argtypes = [vp, c_int32]
    lib.ipl_ib_stage.restype = c_int32
    lib.ipl_ib_body_state.argtypes = [vp, _DP]
    lib.ipl_ib_body_state.restype = c_int32
    lib.ipl_ib_read_markers.argtypes = [vp, c_int32, _DP, c_size_t]
    lib.ipl_ib_read_markers.restype = c_int32
    lib.ipl_batch_create.argtypes = [POINTER(vp)]
    lib.ipl_batch_create.restype = c_int32
    lib.ipl_batch_add.argtypes = [vp, vp]
    lib.ipl_batch_add.restype = c_int32
    lib.ipl_batch_remove.argtypes = [vp, vp]
    lib.ipl_batch_remove.restype = c_int32
    lib.ipl_batch_step.argtypes = [vp, POINTER(c_int32), c_int32]
    lib.ipl_batch_step.restype = c_int32
    lib.ipl_batch_set_launch.argtypes = [vp, c_int32]
    lib.ipl_batch_set_launch.restype = c_int32
    lib.ipl_batch_set_max_lines.argtypes = [vp, c_int64]
    lib.ipl_batch_set_max_lines.restype = c_int32
    lib.ipl_batch_counters.argtypes = [vp, POINTER(c_int64)]
    lib.ipl_batch_counters.restype = c_int32
    lib.ipl_batch_destroy.argtypes = [vp]
    lib.ipl_batch_destroy.restype = c_int32
    lib.ipl_batch_last_error_message.argtypes = [vp]
    lib.ipl_batch_last_error_message.restype = c_char_p
    lib.ipl_estimate_device_bytes.argtypes = [POINTER(IplConfig), POINTER(c_int64)]
    lib.ipl_estimate_device_bytes.restype = c_int32
    lib.ipl_estimate_ib_bytes.argtypes = [POINTER(IplConfig), POINTER(IplIbConfig), POINTER(c_int64)]
    lib.ipl_estimate_ib_bytes.restype = c_int32
    lib.ipl_set_graph.argtypes = [vp, c_int32]
    lib.ipl_set_graph.restype = c_int32
    lib.ipl_prepare.argtypes = [vp]
    lib.ipl_prepare.restype = c_int32
    lib.ipl_set_shared.argtypes = [vp, c_int32]
    lib.ipl_set_shared.restype = c_int32
    lib.ipl_graph_captures.argtypes = [vp, POINTER(c_int32)]
    lib.ipl_graph_captures.restype = c_int32
    lib.ipl_set_open_boundary.argtypes = [vp, POINTER(c_int32), c_int64, _DP, _DP, _DP, _DP]
    lib.ipl_set_open_boundary.restype = c_int32
    lib.ipl_hydrostatic_pressure.argtypes = [POINTER(IplConfig), c_int32, c_double, _DP, c_int64]
    lib.ipl_hydrostatic_pressure.restype = c_int32
    lib.ipl_wall_cells.argtypes = [POINTER(IplConfig), c_int32, POINTER(c_int32), POINTER(c_int64)]
    lib.ipl_wall_cells.restype = c_int32
    lib.ipl_write_cd_csv.argtypes = [c_char_p, _DP, _DP, c_int64]
    lib.ipl_write_cd_csv.restype = c_int32
    lib.ipl_change_dt.argtypes = [vp, c_double]
    lib.ipl_change_dt.restype = c_int32
    for name in ("ipl_write_fields_csv", "ipl_write_markers_csv", "ipl_write_vtu"):
        getattr(lib, name).argtypes = [vp, c_char_p]
        getattr(lib, name).restype = c_int32
    _LIBS[path] = lib
    return lib


class IplError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ipl error {code}: {msg}")
        self.code = code
