"""The benchmark workloads at their full sizes in FP32 storage against FP64 storage on the GPU.

The FP64 fused path is the oracle's restatement to <= 1e-12 (the parity tests of the reduced
cases, and the lock exchange and solitary wave at full size over 40 / 100 steps); at these sizes
and step counts the CPU oracle would take hours, so FP32 is measured against the FP64 GPU path.
Tolerances are the targets of FP32 storage: relative L2 <= 1e-3 on the velocity vector and
salinity, drag within 0.5 %.  Measured: lock exchange 5000 steps, velocity 7.3e-5, S 2.7e-5;
solitary wave 1000 steps, velocity 7.6e-5 (phi 1.8e-3, see its test); falling sphere 300 steps,
velocity 2.6e-6, body velocity 2.3e-6, body force 5.5e-6."""
import numpy as np
import pytest

from ip_cu_lbm_amd import F32, F64, PATH_FUSED, Simulation, cases
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _run(case, steps, fields, body=False):
    with Simulation(case) as sim:
        left = steps
        while left > 0:
            k = min(1000, left)
            sim.step(k)
            left -= k
        out = {f: sim.read(f) for f in fields}
        if body:
            out["body"] = sim.body_state()
    return out


def test_lock_exchange_full_size_fp32_vs_fp64_5000_steps():
    """512 x 32 x 64 D3Q19 + D3Q7, z stretched by 1.02: the configuration's full 5000 steps."""
    fields = ("p", "u", "v", "w", "S")
    a = _run(cases.lock_exchange(storage=F32, path=PATH_FUSED), 5000, fields)
    b = _run(cases.lock_exchange(storage=F64, path=PATH_FUSED), 5000, fields)
    vel = rel_l2(np.stack([a[c] for c in "uvw"]), np.stack([b[c] for c in "uvw"]))
    errs = {"vel": vel, "S": rel_l2(a["S"], b["S"]), "p": rel_l2(a["p"], b["p"])}
    print("lock exchange FP32 vs FP64, 5000 steps:", errs, flush=True)
    assert all(np.all(np.isfinite(a[f])) for f in fields)
    assert errs["vel"] <= 1e-3 and errs["S"] <= 1e-3, errs


def test_solitary_wave_full_size_fp32_vs_fp64_1000_steps():
    """1024 x 256 D2Q9 + D2Q9 Allen-Cahn phase field.  The velocity is held to the storage target.
    phi is stated, not a target: the flow amplifies round-off near the interface (one ulp of one
    phi value in the FP64 oracle grows to 3.7e-4 in phi after 1000 steps), and FP32 rounding,
    about 1e-7 of every value every step, moves the interface's cells by more: measured 1.8e-3
    after 1000 steps, held here below 5e-3."""
    fields = ("p", "u", "w", "phi")
    a = _run(cases.solitary_wave(storage=F32, path=PATH_FUSED), 1000, fields)
    b = _run(cases.solitary_wave(storage=F64, path=PATH_FUSED), 1000, fields)
    vel = rel_l2(np.stack([a["u"], a["w"]]), np.stack([b["u"], b["w"]]))
    errs = {"vel": vel, "phi": rel_l2(a["phi"], b["phi"]), "p": rel_l2(a["p"], b["p"])}
    print("solitary wave FP32 vs FP64, 1000 steps:", errs, flush=True)
    assert errs["vel"] <= 1e-3 and errs["phi"] <= 5e-3, errs


def test_falling_sphere_full_size_fp32_vs_fp64():
    """192 x 192 x 384 D3Q19, 1800 markers, multi-direct forcing with 3 passes and the 4-point
    Peskin kernel: velocity field, body velocity and hydrodynamic force after 300 steps."""
    fields = ("u", "v", "w")
    a = _run(cases.falling_sphere(storage=F32, path=PATH_FUSED), 300, fields, body=True)
    b = _run(cases.falling_sphere(storage=F64, path=PATH_FUSED), 300, fields, body=True)
    vel = rel_l2(np.stack([a[c] for c in "uvw"]), np.stack([b[c] for c in "uvw"]))
    bv = rel_l2(a["body"]["vel"], b["body"]["vel"])
    bf = rel_l2(a["body"]["force"], b["body"]["force"])
    print("falling sphere FP32 vs FP64, 300 steps:", {"vel": vel, "body_vel": bv, "body_force": bf}, flush=True)
    assert vel <= 1e-3 and bv <= 1e-3 and bf <= 5e-3
