"""Open (inflow / outflow) boundary on the GPU vs the FP64 oracle (set_velocity1 +
set_wall_boundary1, Karman_vortex_2D/Karman_vortexxz.cu:38-57 and :391-430): the inlet and outlet
columns are reset to the equilibrium of their prescribed (p, u) after every step.  FP64 storage,
<= 1e-12 relative L2 after N steps, through the C ABI (graph replay on)."""
import numpy as np
import pytest

from ip_cu_lbm_amd import F32, F64, PATH_FUSED, Simulation, _abi, cases, wall_cells
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _run(case, cells, vals, steps, oracle_lib, fields=("p", "u", "w")):
    with Simulation(case) as sim:
        sim.set_open_boundary(cells, *vals)
        sim.step(steps)
        g = {f: sim.read(f) for f in fields}
        gb = sim.body_state() if case.body is not None else None
    o = oracle_lib.Oracle(case)
    o.set_open_boundary(cells, *vals)
    o.step(steps)
    r = {f: o.read(f) for f in fields}
    ob = o.body_state() if case.body is not None else None
    o.close()
    return g, r, gb, ob


def test_open_channel_stretched_fp64(oracle_lib):
    """x-stretched channel (flux axis x), slip walls, inlet U, outlet held at its initial state,
    a Taylor-Green disturbance inside."""
    U = 0.02
    case = cases.lattice_taylor_green(nx=48, nz=20, path=PATH_FUSED).replace(periodic=(0, 0, 0), wall_bc=_abi.BC_SLIP)
    case.init = dict(case.init, u=case.init["u"] + U)
    cells = np.concatenate([wall_cells(case, 3), wall_cells(case, 1)])
    nin = wall_cells(case, 3).size
    u_b = np.concatenate([np.full(nin, U), case.init["u"][wall_cells(case, 1)]])
    w_b = np.concatenate([np.zeros(nin), case.init["w"][wall_cells(case, 1)] if "w" in case.init else np.zeros(cells.size - nin)])
    p_b = np.asarray(case.init["p"])[cells]
    g, r, _, _ = _run(case, cells, (p_b, u_b, 0.0, w_b), 150, oracle_lib)
    errs = {f: rel_l2(g[f], r[f]) for f in g}
    print("open channel", errs, flush=True)
    assert max(errs.values()) <= 1e-12


def test_karman_cylinder_fp64(oracle_lib):
    """A fixed cylinder (stress method) in a channel with a uniform inflow: the Karman driver's
    set-up at small size."""
    U = 0.03
    c = cases.cylinder_2d(n=(80, 40), markers=120, storage=F64, path=PATH_FUSED, gravity=0.0)
    c = c.replace(wall_bc=_abi.BC_SLIP)
    c.body.pos = (25.3, 0.5, 20.2)
    c.body.moving = 0
    n = c.n_cells
    c.init = {"u": np.full(n, U), "p": cases.uniform(4, n, -1e-5, 1e-5)}
    cells = np.concatenate([wall_cells(c, 3), wall_cells(c, 1)])
    g, r, gb, ob = _run(c, cells, (0.0, U, 0.0, 0.0), 120, oracle_lib)
    errs = {f: rel_l2(g[f], r[f]) for f in g}
    errs["force"] = rel_l2(gb["force"], ob["force"])
    print("karman", errs, gb["force"], flush=True)
    assert max(errs.values()) <= 1e-12
    assert gb["force"][0] > 0  # drag along the stream


def test_open_boundary_fp32_and_removal(oracle_lib):
    """FP32 storage within 1e-5; removing the boundary (n = 0) restores plain stepping."""
    U = 0.02
    case = cases.lattice_taylor_green(nx=40, nz=16, path=PATH_FUSED, storage=F32).replace(periodic=(0, 0, 0),
                                                                                        wall_bc=_abi.BC_SLIP)
    case.init = dict(case.init, u=case.init["u"] + U)
    cells = wall_cells(case, 3)
    g, r, _, _ = _run(case, cells, (0.0, U, 0.0, 0.0), 100, oracle_lib, fields=("u", "w"))
    assert max(rel_l2(g[f], r[f]) for f in g) <= 1e-5
    with Simulation(case) as a, Simulation(case) as b:
        a.set_open_boundary(cells, 0.0, U, 0.0, 0.0)
        a.set_open_boundary([], 0.0, 0.0, 0.0, 0.0)
        a.step(30)
        b.step(30)
        assert np.array_equal(a.read("u"), b.read("u"))


@pytest.mark.parametrize("make", [
    lambda st: cases.lock_exchange(nx=32, ny=6, nz=16, storage=st, path=PATH_FUSED).replace(periodic=(0, 0, 0)),
    lambda st: cases.lattice_taylor_green(nx=48, nz=24, storage=st, path=PATH_FUSED).replace(periodic=(0, 0, 0)),
    lambda st: cases.lock_exchange(nx=24, ny=6, nz=12, storage=st, path=PATH_FUSED).replace(
        periodic=(0, 0, 0), flux_axes=_abi.FLUX_NONE, wall_bc=_abi.BC_BOUNCE_BACK),
], ids=["le_stretched_z", "tg_2d_stretched_x", "le_uniform_bb"])
@pytest.mark.parametrize("storage", [F64, F32])
def test_wall_function_fp64(oracle_lib, make, storage):
    """Spalding wall shear on the cells next to every wall (wall_function, collision.cu:28-55, the
    driver's six calls at LockExchange.cu:214-219) vs the oracle: FP64 <= 1e-12, FP32 <= 1e-5."""
    case = make(storage).replace(wall_function=1, wall_nu=0.005)
    X, _, Z = case.coords()
    case.init = dict(case.init, u=np.asarray(case.init.get("u", 0.0)) + 0.02 * np.cos(Z.ravel() / 3.0))
    fields = ("p", "u", "v", "w") + (("S",) if case.scalar == _abi.SCALAR_SALINITY else ())
    with Simulation(case) as sim:
        sim.step(80)
        g = {f: sim.read(f) for f in fields}
    o = oracle_lib.Oracle(case)
    o.step(80)
    r = {f: o.read(f) for f in fields}
    o.close()
    vel = rel_l2(np.stack([g[c] for c in "uvw"]), np.stack([r[c] for c in "uvw"]))
    errs = {"vel": vel, "p": rel_l2(g["p"], r["p"])}
    if "S" in g:
        errs["S"] = rel_l2(g["S"], r["S"])
    print(case.name, storage, errs, flush=True)
    tol = 1e-12 if storage == F64 else 1e-5
    assert errs["vel"] <= tol and errs.get("S", 0.0) <= tol
    if storage == F64:
        assert errs["p"] <= tol


@pytest.mark.parametrize("make", [
    lambda: cases.lock_exchange(nx=24, ny=6, nz=12, path=PATH_FUSED),
    lambda: cases.lattice_taylor_green(nx=48, nz=24, path=PATH_FUSED),
], ids=["lock_exchange", "taylor_green"])
def test_change_dt_fp64(oracle_lib, make):
    """reset_items_dt + reset_pressure (basic.cu:100-110, 262-273) as the driver uses them
    (LockExchange.cu:201-211): a high-resolution start (dt / 2), then back to the original dt."""
    case = make()
    fields = ("p", "u", "v", "w")
    with Simulation(case) as sim:
        o = oracle_lib.Oracle(case)
        for k, alpha in ((20, 2.0), (30, 0.5), (25, None)):
            sim.step(k)
            o.step(k)
            if alpha:
                sim.change_dt(alpha)
                o.change_dt(alpha)
        g = {f: sim.read(f) for f in fields}
        r = {f: o.read(f) for f in fields}
        o.close()
    errs = {"vel": rel_l2(np.stack([g[c] for c in "uvw"]), np.stack([r[c] for c in "uvw"])), "p": rel_l2(g["p"], r["p"])}
    print(case.name, errs, flush=True)
    assert max(errs.values()) <= 1e-12
