"""Runs cases through the CPU emulation build of the library and the FP64 oracle and prints the
relative L2 errors as JSON (test infrastructure; tests/test_emu.py runs it in a child process
with the AddressSanitizer runtime preloaded).

    python tests/emu/emu_check.py LIBRARY CASE [CASE ...]

Each CASE is name:storage:steps, e.g. lock_exchange:f16:150; the case "batch" (batch:mix:0) runs
a batch of mixed jobs with uneven step counts and compares every member with the same job stepped
alone (bit-exact)."""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from ip_cu_lbm_amd import F16, F32, F64, PATH_FUSED, PATH_PLAIN, Batch, Simulation, _abi, cases  # noqa: E402

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
    if name == "phase_field_3d":
        return cases.phase_field_3d(storage=storage, path=PATH_FUSED), ("p", "u", "v", "w", "phi")
    if name in ("sphere_mdf", "sphere_stress"):  # a moving sphere: every IB kernel, markers re-ranked each step
        mdf = name == "sphere_mdf"
        n = (16, 14, 18)
        c = cases.falling_sphere(n=n, diameter=6.0, method=_abi.IB_DIRECT if mdf else _abi.IB_STRESS,
                                 kernel=_abi.IB_KERNEL_PESKIN4 if mdf else _abi.IB_KERNEL_REFERENCE, iterations=3 if mdf else 1,
                                 storage=storage, path=PATH_FUSED, marker_set="2", gravity=2e-3, tau=0.7,
                                 density_ratio=8.0 if mdf else 1.5)
        c.body.pos = (n[0] / 2 + 0.3, n[1] / 2 - 0.2, n[2] / 2 + 0.4)
        X, Y, Z = c.coords()
        c.init = {"u": 0.02 * np.sin(2 * np.pi * Z / n[2]), "p": cases.uniform(5, c.n_cells, -1e-4, 1e-4)}
        return c, ("p", "u", "v", "w")
    if name == "box_d3q27":
        return cases.lock_exchange(nx=12, ny=8, nz=10, ratio=1.0, storage=storage, scalar_lattice=0, path=PATH_FUSED).replace(
            lattice=27, scalar=0, gravity=0.0, flux_axes=0), ("p", "u", "v", "w")
    raise KeyError(name)


def batch_check(lib):
    """Members of two launch groups of the batched kernel (D2Q9 FP32 at three sizes, D3Q19+D3Q7
    FP32 at two, D2Q9 FP64), an FP16 member and an immersed-boundary member (both stepped on their
    own streams), uneven counts with zeros and odd counts: each member against itself alone."""
    jobs = [cases.lattice_taylor_green(nx=24, nz=16, storage=F32, path=PATH_FUSED),
            cases.lattice_taylor_green(nx=32, nz=12, storage=F32, path=PATH_FUSED),
            cases.lock_exchange(nx=16, ny=4, nz=8, storage=F32, path=PATH_FUSED),
            cases.lattice_taylor_green(nx=40, nz=20, storage=F32, path=PATH_FUSED),
            cases.lock_exchange(nx=12, ny=6, nz=10, storage=F32, path=PATH_FUSED),
            cases.lattice_taylor_green(nx=20, nz=12, storage=F64, path=PATH_FUSED),
            cases.lattice_taylor_green(nx=16, nz=16, storage=F16, path=PATH_FUSED),
            make("cylinder", F32)[0]]
    counts = [[3, 2, 4], [1, 1, 1], [4, 0, 3], [0, 5, 2], [2, 2, 1], [3, 4, 0], [2, 1, 2], [1, 2, 1]]
    fields = ("p", "u", "w")
    alone = []
    for c, k in zip(jobs, counts):
        with Simulation(c, lib) as s:
            s.step(sum(k))
            alone.append([s.read(f) for f in fields])
    sims = [Simulation(c, lib) for c in jobs]
    try:
        with Batch(lib) as b:
            for s in sims:
                b.add(s)
            for r in range(3):
                b.step([k[r] for k in counts])
            ctr = b.counters()
        got = [[s.read(f) for f in fields] for s in sims]
    finally:
        for s in sims:
            s.close()
    same = [all(np.array_equal(x, y) for x, y in zip(g, a)) for g, a in zip(got, alone)]
    return {"identical": same, "finite": bool(all(np.all(np.isfinite(x)) for g in got for x in g)), "path": 2,
            "counters": ctr, "batched_expected": sum(sum(k) for k in counts[:6]), "own_expected": sum(sum(k) for k in counts[6:])}


def rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a))


def main(argv):
    lib = argv[0]
    out = {}
    for spec in argv[1:]:
        name, st, steps = spec.split(":")
        if name == "batch":
            out[spec] = batch_check(lib)
            print(f"[emu] {spec}: {out[spec]}", file=sys.stderr, flush=True)
            continue
        case, fields = make(name, STORAGE[st])
        if "S" not in fields and case.init.get("S") is not None and case.scalar == 0:
            case.init = {k: v for k, v in case.init.items() if k != "S"}
        with Simulation(case, lib) as sim:
            sim.step(int(steps))
            got = {f: sim.read(f) for f in fields}
            path = sim.path
            counts = (sim.n_calc, sim.n_ghost)
            if case.body is not None:
                gbody = (sim.body_state()["vel"], sim.markers("force"), sim.markers("node"))
        o = oracle.Oracle(case.replace(storage=F64))
        try:
            same_counts = counts == (o.n_calc, o.n_ghost)
            o.step(int(steps))
            want = {f: o.read(f) for f in fields}
            if case.body is not None:
                obody = (o.body_state()["vel"], o.markers("force"), o.markers("node"))
        finally:
            o.close()
        vel = [f for f in ("u", "v", "w") if f in fields]
        errs = {f: rel(got[f], want[f]) for f in fields}
        errs["vel"] = rel(np.concatenate([got[f] for f in vel]), np.concatenate([want[f] for f in vel]))
        errs["finite"] = bool(all(np.all(np.isfinite(got[f])) for f in fields))
        errs["path"] = path
        errs["counts"] = same_counts  # calc and ghost cells as the reference's builder counts them
        if case.body is not None:
            errs["body_vel"] = rel(gbody[0], obody[0])
            errs["Gw"] = rel(gbody[1], obody[1])
            errs["nodes"] = bool(np.array_equal(gbody[2], obody[2]))
        out[spec] = errs
        print(f"[emu] {spec}: {errs}", file=sys.stderr, flush=True)
    print(json.dumps(out))


if __name__ == "__main__":
    main(sys.argv[1:])
