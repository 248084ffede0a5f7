"""The CPU emulation build of the HIP sources (tests/emu/), under AddressSanitizer, against the
oracle (no GPU).

The emulation runs every kernel of the library on the host: work-items of a synchronising
workgroup are fibers that yield at __syncthreads, device memory is host memory, and every
population or field access of the fused step is checked against the allocation that holds its
row.  An out-of-bounds index anywhere therefore aborts the child process instead of faulting a
GPU.  The FP16-storage instantiations, which the GPU library does not carry, are built here so
that their host and kernel code stays examined.

Tolerances (relative L2 against the FP64 oracle): FP64 1e-12 on every field; FP32 1e-5 on the
velocity vector; FP16 storage (deviation form, scaled by 2^13) 2e-3 on the lock exchange's
velocity after 150 steps and 5e-3 on the solitary wave's after 60: measured 1.1e-3 and 3.0e-3,
recorded here so that a change of the FP16 arithmetic shows.  The batch case steps a mix of jobs
in batched launches (k_fused_step) and requires each to end bit-identical to itself alone."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests", "emu"))

import build_emu  # noqa: E402

CASES = ["lock_exchange:f64:60", "lock_exchange:f32:60", "lock_exchange:f16:150", "solitary_wave:f64:60",
         "solitary_wave:f16:60", "lock_exchange_plain:f64:30", "cylinder:f64:20", "box_d3q27:f64:20",
         "sphere_mdf:f64:12", "sphere_stress:f64:12", "phase_field_3d:f64:40", "batch:mix:0"]


@pytest.fixture(scope="module")
def emu_results():
    lib = build_emu.build(asan=True)
    asan = subprocess.run(["g++", "-print-file-name=libasan.so"], check=True, capture_output=True, text=True).stdout.strip()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", OMP_NUM_THREADS="4")
    env["LD_PRELOAD"] = (asan + " " + env["LD_PRELOAD"]) if env.get("LD_PRELOAD") else asan
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "emu", "emu_check.py"), lib] + CASES,
                       capture_output=True, text=True, env=env, timeout=900)
    assert r.returncode == 0, r.stderr[-6000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "[emu] out of bounds" not in r.stderr, r.stderr[-6000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


@pytest.mark.parametrize("spec", CASES)
def test_emulation_matches_oracle(emu_results, spec):
    e = emu_results[spec]
    assert e["finite"], spec
    st = spec.split(":")[1]
    if spec.startswith("batch"):  # batched launches: each member bit-identical to itself alone
        assert all(e["identical"]), e
        c = e["counters"]
        assert c["batched_steps"] == e["batched_expected"] and c["own_steps"] == e["own_expected"], e
        assert 0 < c["launches"] <= c["batched_steps"], e
        return
    assert e["counts"], spec
    if "nodes" in e:  # immersed body: nearest cells equal, marker forces within 1e-10 (see the GPU tests)
        assert e["nodes"], spec
        assert e.pop("Gw") <= 1e-10, (spec, e)
    if st == "f64":
        for f, v in e.items():
            if f in ("finite", "path", "v", "counts", "nodes"):  # v is near zero in these flows: its own norm is round-off
                continue
            assert v <= 1e-12, (spec, f, v)
    elif st == "f32":
        assert e["vel"] <= 1e-5, (spec, e)
    else:
        assert e["vel"] <= (2e-3 if spec.startswith("lock") else 5e-3), (spec, e)
    assert e["path"] == (1 if "plain" in spec else 2)
