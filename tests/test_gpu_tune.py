"""Line axis of the fused step on uniform grids: every forced axis matches the oracle, a
stretched grid refuses an axis other than its stretched one, and the tuner picks an admissible
axis whose run equals the library's own choice."""
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
    assert res["line_axis"] in (1, 2, 3) and set(res["ms_per_step"]) <= {"x", "y", "z"}
    assert all(t > 0 for t in res["ms_per_step"].values()) and res["auto_ms_per_step"] > 0
    tuned = tune.apply(case, cache)
    assert tuned.line_axis == res["line_axis"]
    with Simulation(tuned) as a, Simulation(case) as b:
        a.step(30)
        b.step(30)
        for f in ("u", "w", "S"):
            x, y = a.read(f), b.read(f)
            assert np.linalg.norm(x - y) <= 1e-5 * np.linalg.norm(y), f
