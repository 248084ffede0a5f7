#!/usr/bin/env python3
"""Benchmark of the interpolation-based lattice Boltzmann step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload: the stratified lock exchange, D3Q19 MRT flow + D3Q7 salinity with Boussinesq
buoyancy, 512 x 32 x 64 cells, z stretched by 1.02 toward mid-depth (ip_cu_lbm_amd.cases.
lock_exchange), synthetic seeded initial fields.  The headline value is the MLUPS of the fused
step with FP32 populations (FP32 arithmetic); the plain FP64 path (one kernel per reference
stage) and the fused FP64 path are measured beside it.  Inputs are resident in HBM before the
timed region; timing uses HIP events on the library's stream, K steps between two
synchronisations.

The roofline object is for the dominant kernel (the fused step): algorithmic bytes per launch =
2 x (19 + 7) populations x 4 B per cell (each population read once and written once) x cells,
divided by the kernel's mean launch duration over the timed region (HIP events around the K
timed steps, which launch only this kernel on this workload; the other kernels' time from the
per-launch profile, zero here, is subtracted); peak = 8 TB/s HBM.
"traffic" is the measured HBM bytes per launch from rocprofv3 FETCH_SIZE / WRITE_SIZE when a
counter summary for this workload exists under profiles/ (see --pmc-summary), else null.

The CPU baseline is the FP64 oracle (a C++/OpenMP restatement of the reference kernels) on a
bounded number of steps of the same workload, on the host's cores.

Beside the headline the line carries: the fused FP64 and plain FP64/FP32 paths on the same
workload, the falling sphere with its immersed boundary, the phase-field solitary wave and the
stretched Taylor-Green vortex, the falling-sphere driver's grid refined in x and y (two-axis
corner scheme, plain path, FP32 and FP64), the line-axis tuner's timings and the serving mix
(p50/p99 latency, aggregate MLUPS).  --precision runs the lock exchange and the solitary wave to their full step
counts in FP32 and FP64 (and, with --precision-oracle, against the oracle), --sphere-precision
the falling sphere.

Only one GPU is used; --gpus above 1 exits with a message.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBPS = 8000.0
WORKLOAD = "lock_exchange_512x32x64_d3q19_d3q7"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", f"traffic_{WORKLOAD}_fp32.json")
SPHERE_WORKLOAD = "falling_sphere_192x192x384_d3q19"
SPHERE_TRAFFIC_FILE = os.path.join(ROOT, "profiles", f"traffic_{SPHERE_WORKLOAD}_fp32.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measure(case, steps, warmup, profile_steps=0, copy_reference=False):
    """MLUPS of `steps` steps after `warmup` steps, plus per-launch profile.  The step graphs of
    both buffer parities are captured before the timed region (ipl_prepare), so it holds steps
    only.  copy_reference: afterwards the flow populations are written back through ipl_write,
    whose k_store_pops launches (an FP64 row read, an FP32 row written, per population) are the
    copy of known size that the traffic counters are scaled by."""
    from ip_cu_lbm_amd import Simulation
    with Simulation(case) as sim:
        sim.step(warmup)
        sim.prepare()
        t0 = time.perf_counter()
        ms = sim.step_timed(steps)
        wall = time.perf_counter() - t0
        prof = sim.profile_steps(profile_steps) if profile_steps else None
        n = sim.n_calc
        path = sim.path
        sums = sim.global_sums()
        if copy_reference:
            sim.write("f", sim.read("f"))
    return {"mlups": n * steps / (ms * 1e3), "ms_per_step": ms / steps, "wall_s": wall, "n_cells": n,
            "path": path, "profile": prof, "sums": sums.tolist()}


def measure_sphere(steps, warmup, copy_reference=False):
    """The falling-sphere workload (192 x 192 x 384 D3Q19 FP32, 1800 Fibonacci markers, multi-
    direct forcing with 3 passes and the 4-point Peskin kernel, FP64 body): MLUPS of the whole
    step, the fused kernel's time and bandwidth (2 x 19 x 4 B per cell), the IB sub-step's share.
    copy_reference: as in measure(), the populations are written back afterwards (the copy the
    traffic counters are scaled by)."""
    from ip_cu_lbm_amd import Simulation, cases
    case = cases.falling_sphere()
    with Simulation(case) as sim:
        sim.step(warmup)
        sim.prepare()
        ms = sim.step_timed(steps)
        prof = sim.profile_steps(min(steps, 50))
        body = sim.body_state()
        n = sim.n_calc
        if copy_reference:
            sim.write("f", sim.read("f"))
    step_us = ms / steps * 1e3
    kern_us = prof[0] * 1e3
    return {"workload": "falling_sphere_192x192x384_d3q19_ibm_mdf3_peskin4", "cells": n,
            "markers": len(case.body.markers), "mlups": round(n * steps / (ms * 1e3), 1),
            "step_us": round(step_us, 1), "kernel_us": round(kern_us, 1),
            "kernel_gbps": round(2 * 19 * 4 * n / (kern_us * 1e-6) / 1e9, 1),
            "ib_us": round(step_us - kern_us, 1), "steps": warmup + steps,
            "body_vel_z": float(body["vel"][2]), "body_force_z": float(body["force"][2])}


def measure_sphere_precision(steps=20000, every=1000):
    """The falling-sphere workload run to `steps` in FP32 and in FP64 storage (the FP64 GPU path
    matches the FP64 oracle to 1e-12 on the reduced cases; the oracle itself cannot run 2.8e11
    cell updates here).  Every `every` steps: the body velocity and the hydrodynamic drag along z
    (the body force minus the constant (1 - rho_f/rho_B) m g).  Reported at the terminal state
    (the step of largest |v|, before the sphere reaches the bottom wall): settling velocity, drag,
    particle Reynolds number |v| D / nu with nu = (tau - 1/2) c^2 dt / 3, and the FP32 errors
    against FP64; plus the FP32 velocity-field error at the end."""
    import numpy as np
    from ip_cu_lbm_amd import F32, F64, Simulation, cases
    out = {}
    for name, st in (("fp64", F64), ("fp32", F32)):
        case = cases.falling_sphere(storage=st)
        bdy = case.body
        w_net = (1.0 - bdy.rho_f / bdy.density) * bdy.mass * bdy.gravity[2]
        hist = []
        with Simulation(case) as sim:
            for k in range(0, steps, every):
                sim.step(every)
                b = sim.body_state()
                hist.append((k + every, float(b["vel"][2]), float(b["force"][2]) - w_net))
                log(f"[sphere {name}] step {k + every}: v_z {hist[-1][1]:.6e}, drag_z {hist[-1][2]:.6e}")
            u = np.stack([sim.read(c) for c in "uvw"])
        out[name] = {"hist": hist, "u": u, "case": case}
    c = out["fp64"]["case"]
    nu = (c.tau - 0.5) * (c.dz / c.dt) ** 2 * c.dt / 3.0
    d = float(2 * np.abs(c.body.markers).max())
    h64, h32 = out["fp64"]["hist"], out["fp32"]["hist"]
    t = int(np.argmax([abs(v) for _, v, _ in h64]))
    v64, v32, f64, f32 = h64[t][1], h32[t][1], h64[t][2], h32[t][2]
    du = float(np.linalg.norm(out["fp32"]["u"] - out["fp64"]["u"]) / np.linalg.norm(out["fp64"]["u"]))
    return {"workload": "falling_sphere_192x192x384_d3q19_ibm_mdf3_peskin4", "steps": steps, "terminal_step": h64[t][0],
            "settling_velocity_fp64": v64, "settling_velocity_fp32": v32, "drag_fp64": f64, "drag_fp32": f32,
            "re_p": abs(v64) * d / nu, "settling_rel_fp32_vs_fp64": abs(v32 - v64) / abs(v64),
            "drag_rel_fp32_vs_fp64": abs(f32 - f64) / abs(f64), "velocity_rel_l2_fp32_vs_fp64_at_end": du,
            "history_fp64": h64, "history_fp32": h32}


def _rel(a, b):
    import numpy as np
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def measure_precision(with_oracle=False):
    """Relative L2 errors of every field after the configurations' full step counts:
    - lock exchange 512 x 32 x 64 (D3Q19 + D3Q7 salinity), 5000 steps;
    - phase-field solitary wave 1024 x 256 (D2Q9 + D2Q9 phi), 10000 steps.
    For each: the fused FP32 path against the fused FP64 path, and the fused FP64 path against the
    plain FP64 path (one kernel per reference stage).  Both FP64 GPU paths match the FP64 oracle to
    1e-12 on the reduced cases of the GPU tests; with `with_oracle` the solitary wave's fused FP64
    result is also compared with the oracle itself at full size (OpenMP on the host, minutes)."""
    import numpy as np
    from ip_cu_lbm_amd import F32, F64, PATH_FUSED, PATH_PLAIN, Simulation, cases
    out = {}
    for name, make, steps, fields in (
            ("lock_exchange_512x32x64_d3q19_d3q7", cases.lock_exchange, 5000, ("p", "u", "v", "w", "S")),
            ("solitary_wave_1024x256_d2q9_phase_field", cases.solitary_wave, 10000, ("p", "u", "w", "phi"))):
        res = {}
        for tag, st, path in (("fused_fp64", F64, PATH_FUSED), ("fused_fp32", F32, PATH_FUSED), ("plain_fp64", F64, PATH_PLAIN)):
            with Simulation(make(storage=st, path=path)) as sim:
                left = steps
                while left > 0:
                    k = min(1000, left)
                    sim.step(k)
                    left -= k
                res[tag] = {f: sim.read(f) for f in fields}
            log(f"[precision] {name} {tag}: {steps} steps done")
        for r in res.values():  # the velocity vector (a near-zero component alone is ill-conditioned)
            r["vel"] = np.concatenate([r[c] for c in ("u", "v", "w") if c in r])
        keys = fields + ("vel",)
        entry = {"steps": steps,
                 "fp32_vs_fp64_rel_l2": {f: _rel(res["fused_fp32"][f], res["fused_fp64"][f]) for f in keys},
                 "fused_fp64_vs_plain_fp64_rel_l2": {f: _rel(res["fused_fp64"][f], res["plain_fp64"][f]) for f in keys}}
        if with_oracle and name.startswith("solitary"):
            import oracle
            oracle.build()
            o = oracle.Oracle(make(storage=F64))
            try:
                for k in range(0, steps, 500):
                    o.step(500)
                    log(f"[precision] oracle {name}: step {k + 500}")
                got = {f: o.read(f) for f in fields}
                got["vel"] = np.concatenate([got["u"], got["w"]])
                entry["fused_fp64_vs_oracle_rel_l2"] = {f: _rel(res["fused_fp64"][f], got[f]) for f in keys}
            finally:
                o.close()
            entry["oracle_one_ulp_sensitivity"] = oracle_sensitivity(make(storage=F64), (100, 300, 500, 1000))
        out[name] = entry
        log(f"[precision] {name}: {json.dumps(entry)}")
    if with_oracle:
        out["falling_sphere_192x192x384_d3q19_ibm_mdf3_peskin4"] = sphere_vs_oracle(steps=4)
    return out


def sphere_vs_oracle(steps):
    """The full falling sphere (14,155,776 cells, 1800 markers, multi-direct forcing with the
    Peskin kernel) in FP64 on the fused path against the oracle after `steps` steps: fluid fields,
    marker forces and the body state (the oracle's set-up takes minutes at this size)."""
    import numpy as np
    import oracle
    from ip_cu_lbm_amd import F64, PATH_FUSED, Simulation, cases
    case = cases.falling_sphere(storage=F64, path=PATH_FUSED)
    with Simulation(case) as sim:
        sim.step(steps)
        g = {f: sim.read(f) for f in ("p", "u", "v", "w")}
        gf, gb = sim.markers("force"), sim.body_state()
    oracle.build()
    log("[precision] falling sphere: oracle set-up ...")
    o = oracle.Oracle(case)
    try:
        for k in range(steps):
            o.step(1)
            log(f"[precision] falling sphere: oracle step {k + 1}")
        r = {f: o.read(f) for f in ("p", "u", "v", "w")}
        of, ob = o.markers("force"), o.body_state()
    finally:
        o.close()
    vel = lambda d: np.concatenate([d["u"], d["v"], d["w"]])
    res = {"steps": steps, "p": _rel(g["p"], r["p"]), "vel": _rel(vel(g), vel(r)),
           "marker_force": _rel(gf.ravel(), of.reshape(gf.shape).ravel()),
           "body_force": _rel(gb["force"], ob["force"]), "body_vel": _rel(gb["vel"], ob["vel"])}
    log(f"[precision] falling sphere vs oracle: {res}")
    return res


def oracle_sensitivity(case, checkpoints):
    """The oracle against itself with phi of one interface cell moved by one ulp: the relative L2
    difference of p, the velocity and phi at each checkpoint.  It measures how fast the flow
    amplifies round-off, which bounds how long two correct FP64 implementations that sum in
    different orders can agree."""
    import numpy as np
    import oracle
    pert = case.replace(init=dict(case.init))
    phi = pert.init["phi"].copy()
    i = int(np.argmax(np.abs(phi - 0.5) < 0.4))
    phi[i] = np.nextafter(phi[i], 2.0)
    pert.init["phi"] = phi
    a, b = oracle.Oracle(case), oracle.Oracle(pert)
    out, done = {}, 0
    try:
        for t in checkpoints:
            a.step(t - done)
            b.step(t - done)
            done = t
            va = np.concatenate([a.read("u"), a.read("w")])
            vb = np.concatenate([b.read("u"), b.read("w")])
            out[str(t)] = {"p": _rel(b.read("p"), a.read("p")), "vel": _rel(vb, va), "phi": _rel(b.read("phi"), a.read("phi"))}
            log(f"[precision] one-ulp sensitivity, step {t}: {out[str(t)]}")
    finally:
        a.close()
        b.close()
    return out


def measure_small_configs(steps):
    """The two 2D workloads beside the headline (both fit the 256 MB last-level cache, so their
    bandwidth figure is not an HBM one):
    - phase-field solitary wave, D2Q9 MRT + D2Q9 Allen-Cahn phi, 1024 x 256, z stretched by 1.01
      around the interface, fused path, FP32 and FP64 populations, plain FP64 beside them;
      algorithmic bytes 2 x (9 + 9) x s per cell;
    - Taylor-Green vortex in a stretched channel, D2Q9 MRT, 256 x 128, x stretched by 1.01,
      fused and plain FP64 (2 x 9 x 8 B per cell)."""
    from ip_cu_lbm_amd import F32, F64, PATH_FUSED, PATH_PLAIN, cases
    out = {}
    k = max(200, steps // 2)
    sw32 = measure(cases.solitary_wave(storage=F32, path=PATH_FUSED), k, 50, profile_steps=100)
    sw64 = measure(cases.solitary_wave(storage=F64, path=PATH_FUSED), k, 50, profile_steps=100)
    swp = measure(cases.solitary_wave(storage=F64, path=PATH_PLAIN), max(50, k // 10), 10)
    cells = sw32["n_cells"]
    out["solitary_wave"] = {
        "workload": "solitary_wave_1024x256_d2q9_phase_field", "cells": cells,
        "fused_fp32_mlups": round(sw32["mlups"], 1), "fused_fp64_mlups": round(sw64["mlups"], 1),
        "plain_fp64_mlups": round(swp["mlups"], 1),
        "fused_fp32_kernel_us": round(sw32["profile"][0] * 1e3, 2),
        "fused_fp32_kernel_gbps": round(2 * 18 * 4 * cells / (sw32["profile"][0] * 1e-3) / 1e9, 1),
        "fused_fp32_other_us_per_step": round(sw32["profile"][1] * 1e3, 2),
        "speedup_fused_fp32_vs_plain_fp64": round(sw32["mlups"] / swp["mlups"], 2)}
    tg64 = measure(cases.lattice_taylor_green(storage=F64, path=PATH_FUSED), k, 50, profile_steps=100)
    tgp = measure(cases.lattice_taylor_green(storage=F64, path=PATH_PLAIN), max(50, k // 10), 10)
    out["taylor_green"] = {
        "workload": "taylor_green_256x128_d2q9_stretched_x", "cells": tg64["n_cells"],
        "fused_fp64_mlups": round(tg64["mlups"], 1), "plain_fp64_mlups": round(tgp["mlups"], 1),
        "fused_fp64_kernel_us": round(tg64["profile"][0] * 1e3, 2)}
    return out


def measure_corner_grid(steps=40):
    """The falling-sphere driver's own grid (SphereFall.cu:80-128): 100 coarse columns in x and y,
    refined 2x and 4x around the sphere's path (cases.sphere_fall_divs), 100 uniform cells in z,
    D3Q27 MRT, slip walls.  With x and y both refined the step takes the two-axis corner scheme
    (adv.cu:100-175), which runs on the plain path; FP32 and FP64 populations."""
    import numpy as np
    from ip_cu_lbm_amd import F32, F64, PATH_PLAIN, _abi, cases
    px, wx = cases.subdivided_axis(cases.sphere_fall_divs(100, 0.005), 4.0)
    pz, wz = cases.axis_from_widths(np.ones(100))
    out = {"workload": f"sphere_fall_grid_{len(px)}x{len(px)}x100_d3q27_corner_scheme", "path": "plain"}
    for st, name in ((F32, "fp32"), (F64, "fp64")):
        case = cases.Case(name="sphere_fall_grid", lattice=27, n=(len(px), len(px), 100), pos=[px, px.copy(), pz],
                          width=[wx, wx.copy(), wz], wall_bc=_abi.BC_SLIP, flux_axes=_abi.FLUX_REFERENCE, tau=0.7,
                          path=PATH_PLAIN, storage=st, s_rates=np.full(27, 1.2))
        r = measure(case, steps, 5)
        out["cells"] = r["n_cells"]
        out[f"{name}_mlups"] = round(r["mlups"], 1)
        out[f"{name}_ms_per_step"] = round(r["ms_per_step"], 3)
    return out


def measure_tuner(steps=100):
    """The line-axis autotuner (ip_cu_lbm_amd.tune) on uniform grids: every admissible axis and
    the library's own choice timed over `steps` steps (graph replay, HIP events):
    - the falling sphere 192 x 192 x 384, D3Q19 FP32 with its immersed boundary (whole step);
    - a D3Q19 FP32 box 160 x 96 x 72 (walls, no scalar field);
    - the 2D cylinder 512 x 512, D2Q9 FP32 with 400 markers."""
    from ip_cu_lbm_amd import F32, _abi, cases, tune
    out = []
    box = cases.lock_exchange(nx=160, ny=96, nz=72, ratio=1.0, storage=F32, scalar_lattice=0).replace(
        flux_axes=_abi.FLUX_NONE, scalar=_abi.SCALAR_NONE, gravity=0.0, name="box_160x96x72")
    box.init = {k: v for k, v in box.init.items() if k != "S"}
    for case in (cases.falling_sphere(storage=F32), box,
                 cases.cylinder_2d(n=(512, 512), markers=400, storage=F32)):
        r = tune.tune_line_axis(case, steps=steps, warmup=20)
        best = r["ms_per_step"][tune.AXES[r["line_axis"]]]
        out.append({"workload": case.name, "cells": case.n_cells, "tuned_axis": tune.AXES[r["line_axis"]],
                    "ms_per_step": {a: round(t, 5) for a, t in r["ms_per_step"].items()},
                    "auto_ms_per_step": round(r["auto_ms_per_step"], 5),
                    "tuned_vs_auto": round(r["auto_ms_per_step"] / best, 3)})
    return out


def measure_serving(n_jobs=256):
    """The serving mix: 256 seeded Poisson arrivals (mean gap 20 ms) of D2Q9/D3Q19 FP32 jobs of
    mixed sizes (2D sides 64-512, 3D sides 32-128, 500-5000 steps, ~30 % with salinity, ~10 % with
    an immersed body), admitted under a 200 GB device budget and advanced by continuous batching
    in chunks of 400 steps.  Latency is completion - arrival; aggregate MLUPS is all jobs' cell
    updates over the makespan (first arrival to last completion), busy MLUPS over the time spent
    in batch steps."""
    from ip_cu_lbm_amd import serve
    jobs = serve.serving_mix(n_jobs=n_jobs)
    res = serve.Server(budget_bytes=200e9, chunk=400).run(jobs)
    res["workload"] = f"serving_mix_{n_jobs}_jobs_fp32"
    return {k: (round(v, 4) if isinstance(v, float) else v) for k, v in res.items()}


def cpu_baseline(case, seconds):
    import oracle
    oracle.build()
    o = oracle.Oracle(case)
    t1 = o.time_steps(1)
    n = max(1, min(200, int(seconds / max(t1, 1e-3))))
    t = o.time_steps(n)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(o.n_calc * n / t / 1e6, 3), "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"{n} steps of the {WORKLOAD} case in FP64 (the oracle), {t:.1f} s, "
                      f"OpenMP threads={threads}"}


def read_traffic():
    if not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as fh:
        d = json.load(fh)
    return d.get("hbm_bytes_per_launch")


def pmc_summary(fetch_dir, write_dir, cells, sphere=False):
    """Turns two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE; KiB per dispatch) into the traffic
    file, scaling each counter by its reading on the library's copy of a known-size buffer; of
    the headline, or with `sphere` of the falling sphere's fused kernel (--only-sphere)."""
    import csv
    import glob

    def rows(d):
        out = []
        for fn in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(fn) as fh:
                out += list(csv.DictReader(fh))
        return out

    def mean_for(rs, name_part, counter):
        vals = [float(r["Counter_Value"]) for r in rs if name_part in r["Kernel_Name"] and r["Counter_Name"] == counter]
        return sum(vals) / len(vals) if vals else None

    fr, wr = rows(fetch_dir), rows(write_dir)
    f_main = mean_for(fr, "k_fused_step", "FETCH_SIZE")
    w_main = mean_for(wr, "k_fused_step", "WRITE_SIZE")
    f_copy = mean_for(fr, "k_store_pops", "FETCH_SIZE")   # reads an FP64 row of `cells` values
    w_copy = mean_for(wr, "k_store_pops", "WRITE_SIZE")   # writes an FP32 row of `cells` values
    fs = (cells * 8 / 1024.0) / f_copy if f_copy else 1.0
    ws = (cells * 4 / 1024.0) / w_copy if w_copy else 1.0
    total = (f_main * fs + w_main * ws) * 1024.0
    run = "--only-sphere --steps 10 --warmup 2" if sphere else "--only-headline --steps 20 --warmup 2"
    d = {"workload": SPHERE_WORKLOAD if sphere else WORKLOAD, "fetch_kib": f_main, "write_kib": w_main,
         "fetch_scale": fs, "write_scale": ws, "hbm_bytes_per_launch": total,
         "algorithmic_bytes_per_launch": 2 * (19 if sphere else 26) * 4 * cells,
         "commands": [
             f"rocprofv3 --pmc FETCH_SIZE --output-format csv -d gpurun_out/pmc_fetch -o run -- python bench.py {run}",
             f"rocprofv3 --pmc WRITE_SIZE --output-format csv -d gpurun_out/pmc_write -o run -- python bench.py {run}",
             "python bench.py --pmc-summary gpurun_out/pmc_fetch gpurun_out/pmc_write" + (" --pmc-sphere" if sphere else "")],
         "copy_kernel": f"k_store_pops<float,float,*>: reads an FP64 row and writes an FP32 row of {cells:,} values "
                        f"({cells * 8 // 1024} KiB read, {cells * 4 // 1024} KiB written)"}
    out_file = SPHERE_TRAFFIC_FILE if sphere else TRAFFIC_FILE
    os.makedirs(os.path.dirname(out_file), exist_ok=True)
    with open(out_file, "w") as fh:
        json.dump(d, fh, indent=1)
    print(json.dumps(d))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline sample length")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-compare", action="store_true", help="skip the FP64 plain / fused comparison runs")
    ap.add_argument("--only-headline", action="store_true", help="run just the headline workload (for profilers)")
    ap.add_argument("--only-sphere", action="store_true", help="run just the falling-sphere workload (for profilers)")
    ap.add_argument("--only-tuner", action="store_true", help="run just the line-axis tuner workloads")
    ap.add_argument("--precision", action="store_true",
                    help="FP32 vs FP64 and fused vs plain FP64 field errors of the lock exchange and solitary wave")
    ap.add_argument("--precision-oracle", action="store_true",
                    help="with --precision: also run the oracle on the full solitary wave (minutes)")
    ap.add_argument("--only-small", action="store_true",
                    help="run just the solitary-wave and Taylor-Green workloads (for profilers)")
    ap.add_argument("--sphere-precision", type=int, metavar="STEPS", default=0,
                    help="run the falling sphere STEPS steps in FP32 and FP64 and print their settling velocity and drag")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="write the traffic file from two rocprofv3 --pmc output directories and exit")
    ap.add_argument("--pmc-sphere", action="store_true",
                    help="with --pmc-summary: the runs were of --only-sphere (the falling sphere's traffic file)")
    a = ap.parse_args(argv)

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {a.gpus}, WORLD_SIZE {world}); "
              "independent simulations shard by job, one process per GPU, with no collective in the step.",
              file=sys.stderr)
        return 2

    from ip_cu_lbm_amd import F32, F64, PATH_FUSED, PATH_PLAIN, cases

    if a.pmc_summary:
        if a.pmc_sphere:
            pmc_summary(a.pmc_summary[0], a.pmc_summary[1], 192 * 192 * 384, sphere=True)
        else:
            pmc_summary(a.pmc_summary[0], a.pmc_summary[1], 512 * 32 * 64)
        return 0

    if a.sphere_precision:
        print(json.dumps(measure_sphere_precision(a.sphere_precision)), flush=True)
        return 0

    if a.only_sphere:
        print(json.dumps(measure_sphere(a.steps, a.warmup, copy_reference=True)), flush=True)
        return 0

    if a.only_small:
        print(json.dumps(measure_small_configs(a.steps)), flush=True)
        return 0

    if a.precision:
        print(json.dumps(measure_precision(with_oracle=a.precision_oracle)), flush=True)
        return 0

    if a.only_tuner:
        print(json.dumps(measure_tuner()), flush=True)
        return 0

    head_case = cases.lock_exchange(storage=F32, path=PATH_FUSED)
    log(f"[bench] headline: {WORKLOAD} fused FP32, warmup {a.warmup}, steps {a.steps}")
    head = measure(head_case, a.steps, a.warmup, profile_steps=min(a.steps, 200), copy_reference=a.only_headline)
    cells = head["n_cells"]
    b_alg = 2 * (19 + 7) * 4 * cells
    # The dominant kernel's mean duration over the timed region: the timed steps replay graphs of
    # step launches, and the time of the step's other kernels (none on this workload) measured by
    # the per-launch profile is taken off.  Per-launch HIP events around single launches add their
    # own overhead, so that figure is reported beside it, not used.
    event_ms = head["profile"][0]
    main_ms = head["ms_per_step"] - head["profile"][1]
    achieved = b_alg / (main_ms * 1e-3) / 1e9
    log(f"[bench] fused FP32: {head['mlups']:.0f} MLUPS, kernel {main_ms * 1e3:.1f} us, {achieved:.0f} GB/s")
    # the timed steps against the per-launch profile: a capture or a stall inside the timed region
    # would show as a step far longer than its kernels
    step_vs_kernels = head["ms_per_step"] / (event_ms + head["profile"][1])
    if step_vs_kernels > 1.15:
        log(f"[bench] WARNING: {head['ms_per_step'] * 1e3:.1f} us per timed step against "
            f"{(event_ms + head['profile'][1]) * 1e3:.1f} us of kernels per step (x{step_vs_kernels:.2f})")

    extra = {}
    if not a.no_compare and not a.only_headline:
        k = max(20, a.steps // 20)
        plain64 = measure(cases.lock_exchange(storage=F64, path=PATH_PLAIN), k, 5)
        log(f"[bench] plain FP64: {plain64['mlups']:.0f} MLUPS")
        fused64 = measure(cases.lock_exchange(storage=F64, path=PATH_FUSED), max(100, a.steps // 4), 20,
                          profile_steps=100)
        log(f"[bench] fused FP64: {fused64['mlups']:.0f} MLUPS")
        plain32 = measure(cases.lock_exchange(storage=F32, path=PATH_PLAIN), k, 5)
        log(f"[bench] plain FP32: {plain32['mlups']:.0f} MLUPS")
        sph = measure_sphere(max(100, a.steps // 10), 20)
        log(f"[bench] falling sphere: {sph['mlups']:.0f} MLUPS, fused kernel {sph['kernel_us']:.1f} us, "
            f"IB sub-step {sph['ib_us']:.1f} us per step")
        sw = measure_small_configs(a.steps)
        log(f"[bench] solitary wave: fused FP32 {sw['solitary_wave']['fused_fp32_mlups']:.0f} MLUPS, "
            f"Taylor-Green: fused FP64 {sw['taylor_green']['fused_fp64_mlups']:.0f} MLUPS")
        corner = measure_corner_grid()
        log(f"[bench] sphere-driver grid (corner scheme, plain): FP32 {corner['fp32_mlups']:.0f} MLUPS, "
            f"FP64 {corner['fp64_mlups']:.0f} MLUPS")
        tuned = measure_tuner()
        log("[bench] line-axis tuner: " + "; ".join(f"{t['workload']} {t['tuned_axis']} x{t['tuned_vs_auto']}" for t in tuned))
        srv = measure_serving()
        srv["aggregate_vs_single_large_case"] = round(srv["aggregate_mlups"] / head["mlups"], 3)
        log(f"[bench] serving mix: {srv['jobs']} jobs, p50 {srv['p50_s']:.2f} s, p99 {srv['p99_s']:.2f} s, "
            f"aggregate {srv['aggregate_mlups']:.0f} MLUPS (busy {srv['busy_mlups']:.0f})")
        extra = {
            "fused_fp32_mlups": round(head["mlups"], 1),
            "fused_fp64_mlups": round(fused64["mlups"], 1),
            "plain_fp64_mlups": round(plain64["mlups"], 1),
            "plain_fp32_mlups": round(plain32["mlups"], 1),
            "speedup_fused_fp32_vs_plain_fp64": round(head["mlups"] / plain64["mlups"], 2),
            "fused_fp64_kernel_gbps": round(2 * 26 * 8 * cells / (fused64["profile"][0] * 1e-3) / 1e9, 1),
            "fp16_storage": "disabled: the library refuses IPL_F16 until the fault of its first GPU run is explained",
            "falling_sphere": sph,
            "solitary_wave": sw["solitary_wave"],
            "taylor_green": sw["taylor_green"],
            "sphere_fall_grid": corner,
            "line_axis_tuner": tuned,
            "serving": srv,
        }

    cpu = None
    if not a.no_cpu_baseline and not a.only_headline:
        log("[bench] CPU baseline (FP64 oracle) ...")
        cpu = cpu_baseline(cases.lock_exchange(storage=F64), a.cpu_seconds)
        log(f"[bench] CPU: {cpu['value']} MLUPS on {cpu['cores']} threads")

    line = {
        "metric": "MLUPS & % HBM roofline, FP32/FP16 vs FP64 plain; batched p50/p99 job latency",
        "value": round(head["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(head["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SplitMix64 lock-exchange initial state)",
        "config": {"workload": WORKLOAD, "cells": cells, "lattice": "D3Q19 MRT + D3Q7 salinity",
                   "stretch": "z, ratio 1.02 toward mid-depth", "storage": "fp32", "path": "fused",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBPS, 4), "traffic": read_traffic(),
                     "kernel": "k_fused_step", "kernel_ms": round(main_ms, 5),
                     "kernel_ms_per_launch_events": round(event_ms, 5),
                     "algorithmic_bytes_per_launch": b_alg,
                     "other_ms_per_step": round(head["profile"][1], 5),
                     "step_vs_kernels": round(step_vs_kernels, 3)},
        "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
