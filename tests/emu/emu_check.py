"""Runs cases through the CPU emulation build of the library and the FP64 oracle and prints the
relative L2 errors as JSON (test infrastructure; tests/test_emu.py runs it in a child process
with the AddressSanitizer runtime preloaded).

    python tests/emu/emu_check.py LIBRARY CASE [CASE ...]

Each CASE is name:storage:steps, e.g. lock_exchange:f16:150."""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from ip_cu_lbm_amd import F16, F32, F64, PATH_FUSED, PATH_PLAIN, Simulation, cases  # noqa: E402

STORAGE = {"f64": F64, "f32": F32, "f16": F16}


def make(name, storage):
    """The small workloads of the GPU parity tests, by name."""
    if name == "lock_exchange":
        return cases.lock_exchange(nx=32, ny=6, nz=16, storage=storage, path=PATH_FUSED), ("p", "u", "v", "w", "S")
    if name == "lock_exchange_plain":
        return cases.lock_exchange(nx=16, ny=4, nz=10, storage=storage, path=PATH_PLAIN), ("p", "u", "v", "w", "S")
    if name == "solitary_wave":
        return cases.solitary_wave(nx=96, nz=48, x0=30, sigma=1e-4, storage=storage, path=PATH_FUSED), ("p", "u", "w", "phi")
    if name == "taylor_green":
        return cases.lattice_taylor_green(nx=64, nz=32, storage=storage, path=PATH_FUSED), ("p", "u", "w")
    if name == "cylinder":
        return cases.cylinder_2d(n=(48, 56), markers=140, storage=storage, path=PATH_FUSED, density_ratio=3.0,
                                 gravity=2e-4), ("p", "u", "w")
    if name == "box_d3q27":
        return cases.lock_exchange(nx=12, ny=8, nz=10, ratio=1.0, storage=storage, scalar_lattice=0, path=PATH_FUSED).replace(
            lattice=27, scalar=0, gravity=0.0, flux_axes=0), ("p", "u", "v", "w")
    raise KeyError(name)


def rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a))


def main(argv):
    lib = argv[0]
    out = {}
    for spec in argv[1:]:
        name, st, steps = spec.split(":")
        case, fields = make(name, STORAGE[st])
        if "S" not in fields and case.init.get("S") is not None and case.scalar == 0:
            case.init = {k: v for k, v in case.init.items() if k != "S"}
        with Simulation(case, lib) as sim:
            sim.step(int(steps))
            got = {f: sim.read(f) for f in fields}
            path = sim.path
        o = oracle.Oracle(case.replace(storage=F64))
        try:
            o.step(int(steps))
            want = {f: o.read(f) for f in fields}
        finally:
            o.close()
        vel = [f for f in ("u", "v", "w") if f in fields]
        errs = {f: rel(got[f], want[f]) for f in fields}
        errs["vel"] = rel(np.concatenate([got[f] for f in vel]), np.concatenate([want[f] for f in vel]))
        errs["finite"] = bool(all(np.all(np.isfinite(got[f])) for f in fields))
        errs["path"] = path
        out[spec] = errs
        print(f"[emu] {spec}: {errs}", file=sys.stderr, flush=True)
    print(json.dumps(out))


if __name__ == "__main__":
    main(sys.argv[1:])
