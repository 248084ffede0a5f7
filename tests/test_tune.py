"""Tile autotuner (line axis, workgroup size), host side: shape keys, candidate workgroups, cache
use and persistence (no GPU)."""
from ip_cu_lbm_amd import _abi, cases, tune


def test_shape_key_separates_what_the_speed_depends_on():
    a = cases.lock_exchange(nx=20, ny=12, nz=24)
    assert tune.shape_key(a) == tune.shape_key(a.replace(name="other", tau=0.7))
    for b in (a.replace(n=(20, 12, 25)), a.replace(storage=_abi.F32), a.replace(scalar_lattice=0),
              a.replace(periodic=(1, 0, 0)), a.replace(wall_bc=_abi.BC_BOUNCE_BACK), a.replace(les=1)):
        assert tune.shape_key(b) != tune.shape_key(a)


def test_apply_and_cache_round_trip(tmp_path):
    a = cases.lock_exchange(nx=20, ny=12, nz=24).replace(flux_axes=_abi.FLUX_NONE)
    assert tune.apply(a, {}).line_axis == 0
    cache = {tune.shape_key(a): {"key": tune.shape_key(a), "line_axis": 1, "ms_per_step": {"x": 0.01},
                                 "auto_ms_per_step": 0.02}}
    path = str(tmp_path / "tuned.json")
    tune.save_cache(cache, path)
    back = tune.load_cache(path)
    assert back == cache
    assert tune.apply(a, back).line_axis == 1 and tune.apply(a, back).workgroup == 0
    back[tune.shape_key(a)]["workgroup"] = 256
    assert tune.apply(a, back).workgroup == 256
    assert tune.apply(a.replace(n=(21, 12, 24)), back).line_axis == 0
    assert tune.load_cache(str(tmp_path / "missing.json")) == {}
    # a cached shape is returned without timing anything (no GPU is touched here)
    assert tune.tune_line_axis(a, cache=back) is back[tune.shape_key(a)]


def test_line_axis_reaches_the_config():
    cfg, _ini, _keep = cases.lock_exchange(nx=8, ny=6, nz=10).replace(line_axis=3).to_c()
    assert cfg.line_axis == 3


def test_workgroup_candidates():
    """0 (automatic, one line) first, then the smallest size for each further count of slots."""
    assert tune.workgroups_for(20) == [0, 128, 192, 256, 320, 384, 448, 512]
    assert tune.workgroups_for(100) == [0, 256, 384, 512]
    assert tune.workgroups_for(300) == [0]
    assert tune.workgroups_for(512) == [0]
    cfg, _ini, _keep = cases.lock_exchange(nx=8, ny=6, nz=10).replace(workgroup=128).to_c()
    assert cfg.workgroup == 128
