"""Tile of the fused step: every forced line axis matches the oracle, a stretched grid refuses an
axis other than its stretched one, every workgroup size (one or many lines per workgroup) gives
the same step bit for bit, and the tuner picks an admissible tile whose run equals the library's
own choice."""
import numpy as np
import pytest

from ip_cu_lbm_amd import PATH_FUSED, IplError, Simulation, _abi, cases, tune

pytestmark = pytest.mark.gpu


def _box(**kw):
    return cases.lock_exchange(nx=20, ny=12, nz=24, ratio=1.0, path=PATH_FUSED, **kw).replace(
        periodic=(0, 1, 0), flux_axes=_abi.FLUX_NONE)


@pytest.mark.parametrize("axis", [1, 2, 3], ids=["x", "y", "z"])
def test_forced_line_axis_matches_oracle_fp64(oracle_lib, axis):
    case = _box().replace(line_axis=axis)
    o = oracle_lib.Oracle(case)
    with Simulation(case) as sim:
        assert sim.path == PATH_FUSED
        sim.step(40)
        o.step(40)
        for f in ("p", "u", "w", "S"):
            a, b = sim.read(f), o.read(f)
            err = np.linalg.norm(a - b) / np.linalg.norm(b)
            print(f"line axis {axis} {f}: rel L2 {err:.2e}", flush=True)
            assert err <= 1e-12, (f, err)
    o.close()


def test_stretched_grid_keeps_its_axis():
    case = cases.lock_exchange(nx=20, ny=12, nz=24, path=PATH_FUSED)  # z stretched
    with Simulation(case.replace(line_axis=3)) as sim:
        assert sim.path == PATH_FUSED
    with pytest.raises(IplError):
        Simulation(case.replace(line_axis=1))


def test_tuner_choice_runs_like_the_default():
    case = _box(storage=_abi.F32)
    cache = {}
    res = tune.tune_line_axis(case, steps=20, warmup=5, cache=cache)
    print(res, flush=True)
    assert res["line_axis"] in (1, 2, 3) and res["workgroup"] in tune.workgroups_for(case.n[res["line_axis"] - 1])
    assert {k.split("/")[0] for k in res["ms_per_step"]} <= {"x", "y", "z"}
    assert all(t > 0 for t in res["ms_per_step"].values()) and res["auto_ms_per_step"] > 0
    tuned = tune.apply(case, cache)
    assert tuned.line_axis == res["line_axis"] and tuned.workgroup == res["workgroup"]
    with Simulation(tuned) as a, Simulation(case) as b:
        a.step(30)
        b.step(30)
        for f in ("u", "w", "S"):
            x, y = a.read(f), b.read(f)
            assert np.linalg.norm(x - y) <= 1e-5 * np.linalg.norm(y), f


@pytest.mark.parametrize("storage", [_abi.F64, _abi.F32], ids=["fp64", "fp32"])
def test_workgroup_sizes_step_bit_identical(storage):
    """Lines of 20 cells with 1 to 8 lines per workgroup, a phase-field wave (lines of 64 cells
    along its stretched x) and a stretched lock exchange: the same fields bit for bit."""
    runs = [(_box(storage=storage).replace(line_axis=1), ("p", "u", "v", "w", "S"), (0, 64, 128, 192, 320, 512)),
            (cases.solitary_wave(nx=64, nz=40, x0=20, sigma=1e-4, storage=storage, path=PATH_FUSED),
             ("p", "u", "w", "phi"), (0, 64, 128, 256, 512)),
            (cases.lock_exchange(nx=24, ny=6, nz=40, storage=storage, path=PATH_FUSED), ("p", "u", "w", "S"),
             (0, 64, 128, 256, 512))]
    for case, fields, sizes in runs:
        ref = None
        for w in sizes:
            with Simulation(case.replace(workgroup=w)) as sim:
                assert sim.path == PATH_FUSED
                sim.step(23)
                got = {f: sim.read(f) for f in fields}
            print(case.name, w, flush=True)
            if ref is None:
                ref = got
                continue
            for f in fields:
                assert np.array_equal(got[f], ref[f]), (case.name, w, f)


def test_many_lines_per_workgroup_matches_oracle_fp64(oracle_lib):
    """8 lines of 20 cells in each workgroup of 512 work-items, against the oracle."""
    case = _box().replace(line_axis=1, workgroup=512)
    o = oracle_lib.Oracle(case)
    with Simulation(case) as sim:
        sim.step(30)
        o.step(30)
        for f in ("p", "u", "w", "S"):
            a, b = sim.read(f), o.read(f)
            assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(b), f
    o.close()


def test_workgroup_too_small_for_a_line_is_refused():
    with pytest.raises(IplError):
        Simulation(cases.lock_exchange(nx=20, ny=12, nz=100, ratio=1.0, path=PATH_FUSED).replace(
            periodic=(0, 1, 0), flux_axes=_abi.FLUX_NONE, line_axis=3, workgroup=64))
