"""Seeded sweep of small configurations, GPU (both paths) against the FP64 oracle.

Each seed draws a lattice (D2Q9, D3Q19, D3Q27), a box of a few thousand cells, one stretched flux
axis or none, 3- or 5-point stencils, periodic flags per axis, the wall rule, the collision (MRT
with random rates or SRT), and for D2Q9 / D3Q19 a salinity field (same-Q, or D3Q7) with Boussinesq
buoyancy, on a uniform grid the fused step's line axis, then random initial p, u, v, w (and S).  After 20 steps p, the velocity and S agree with
the oracle to <= 1e-12 (relative L2) on the plain path and, where the configuration is one it
takes, on the fused path.  The sweep covers combinations that no single hand-written case
does (wall rules on periodic-free axes, refined axes next to walls, line axes across walls)."""
import numpy as np
import pytest

from ip_cu_lbm_amd import F32, F64, PATH_FUSED, PATH_PLAIN, Simulation, _abi, cases
from conftest import rel_l2

pytestmark = pytest.mark.gpu

SEEDS = list(range(64))


def _config(seed):
    rng = np.random.default_rng(cases.SEED_BASE + 500 + seed)
    lattice = int(rng.choice([9, 19, 27]))
    if lattice == 9:
        n = (int(rng.integers(8, 40)), 1, int(rng.integers(8, 40)))
        axes = [0, 2]
    else:
        n = tuple(int(v) for v in rng.integers(5, 18, size=3))
        axes = [0, 1, 2]
    flux = int(rng.choice([-1] + axes))  # -1: uniform grid
    ratio = float(rng.uniform(1.02, 1.12))
    pos, wid = [], []
    for a in range(3):
        if a == flux:
            p, w = cases.axis_from_widths(cases.stretched_widths(n[a], ratio))
        else:
            p, w = cases.axis_from_widths(np.ones(n[a]))
        pos.append(p)
        wid.append(w)
    periodic = tuple(int(rng.random() < 0.4) if (a in axes and a != flux) else 0 for a in range(3))
    flux_axes = {-1: _abi.FLUX_NONE, 0: _abi.FLUX_X, 1: _abi.FLUX_Y, 2: _abi.FLUX_Z}[flux]
    collision = _abi.COLL_SRT if rng.random() < 0.25 else _abi.COLL_MRT
    tau = float(rng.uniform(0.55, 0.9))
    kw = dict(name=f"sweep_{seed}", lattice=lattice, n=n, pos=pos, width=wid, periodic=periodic,
              wall_bc=int(rng.choice([_abi.BC_SLIP, _abi.BC_BOUNCE_BACK])), flux_axes=flux_axes,
              stencil=int(rng.choice([3, 5])), collision=collision, s_rates=rng.uniform(1.0, 1.6, size=27),
              tau=tau, storage=F64)
    salinity = lattice in (9, 19) and rng.random() < 0.5
    if salinity:
        kw.update(scalar=_abi.SCALAR_SALINITY, scalar_lattice=7 if (lattice == 19 and rng.random() < 0.5) else 0,
                  taus=float(rng.uniform(0.52, 0.8)), beta_s=1e-3, gravity=float(rng.uniform(0.0, 1e-2)),
                  boussinesq=1)
    if flux == -1:  # a uniform grid: the fused step's lines along any axis
        kw["line_axis"] = int(rng.choice([0, 1, 3] if lattice == 9 else [0, 1, 2, 3]))
    case = cases.Case(**kw)
    X, Y, Z = case.coords()
    init = {"p": cases.uniform(7 + seed, case.n_cells, -1e-3, 1e-3),
            "u": 0.01 * np.sin(X / 2.3 + seed) * np.cos(Z / 1.7),
            "w": 0.01 * np.cos(X / 1.9) * np.sin(Z / 2.1 + seed)}
    if lattice != 9:
        init["v"] = 0.01 * np.sin(Y / 1.3 + seed)
    if salinity:
        init["S"] = 0.5 + 0.3 * np.tanh((X - X.mean()) / 2.0)
    case.init = init
    return case


@pytest.mark.parametrize("seed", SEEDS)
def test_sweep_against_oracle_fp64(oracle_lib, seed):
    base = _config(seed)
    fields = ["p", "u", "w"] + (["v"] if base.lattice != 9 else []) + (["S"] if base.scalar else [])
    o = oracle_lib.Oracle(base)
    o.step(20)
    want = {f: o.read(f) for f in fields}
    o.close()
    ran = []
    for path in (PATH_PLAIN, PATH_FUSED):
        try:
            sim = Simulation(base.replace(path=path))
        except _abi.IplError:
            assert path == PATH_FUSED  # a configuration the fused path does not take
            continue
        with sim:
            sim.step(20)
            got = {f: sim.read(f) for f in fields}
        vel = [f for f in ("u", "v", "w") if f in fields]
        errs = {"p": rel_l2(got["p"], want["p"]),
                "vel": rel_l2(np.concatenate([got[f] for f in vel]), np.concatenate([want[f] for f in vel]))}
        if "S" in fields:
            errs["S"] = rel_l2(got["S"], want["S"])
        print(base.name, base.lattice, base.n, "flux", base.flux_axes, "periodic", base.periodic, "bc", base.wall_bc,
              "coll", base.collision, "scalar", base.scalar, base.scalar_lattice, "path", path, errs, flush=True)
        assert all(np.all(np.isfinite(got[f])) for f in fields)
        for k, e in errs.items():
            assert e <= 1e-12, (path, k, e)
        ran.append(path)
    assert PATH_PLAIN in ran


IB_SEEDS = list(range(32))


def _ib_config(seed):
    """A body in a small box: a sphere (D3Q19 or D3Q27) or a 2D cylinder, the stress method or
    direct forcing (3-point Lagrange or 4-point Peskin, 1-3 passes), moving or fixed, periodic
    flags, and a position anywhere its markers stay inside the box.  Axes whose node weights are
    3-point Lagrange (y and z of the stress method, every axis of direct forcing's reference
    kernel) keep the body a cell from their walls: a ghost node sits on its owner, so beyond a wall
    two of the three positions coincide and the weights are not finite, as in the reference."""
    rng = np.random.default_rng(cases.SEED_BASE + 900 + seed)
    two_d = rng.random() < 0.3
    stress = rng.random() < 0.4
    if stress:
        method, kernel, iters = _abi.IB_STRESS, _abi.IB_KERNEL_REFERENCE, 1
    else:
        method = _abi.IB_DIRECT
        kernel = int(rng.choice([_abi.IB_KERNEL_REFERENCE, _abi.IB_KERNEL_PESKIN4]))
        iters = int(rng.integers(1, 4))
    if two_d:
        n = (int(rng.integers(24, 48)), int(rng.integers(24, 48)))
        c = cases.cylinder_2d(n=n, diameter=float(rng.uniform(6.0, 9.0)), markers=int(rng.integers(60, 160)),
                              storage=F64, path=PATH_FUSED, density_ratio=float(rng.uniform(3.0, 8.0)),
                              gravity=2e-4, method=method)
        dims = (n[0], 1, n[1])
    else:
        dims = tuple(int(v) for v in rng.integers(16, 26, size=3))
        c = cases.falling_sphere(n=dims, diameter=float(rng.uniform(5.0, 8.0)), marker_set=str(int(rng.choice([2, 3]))),
                                 storage=F64, path=PATH_FUSED, tau=float(rng.uniform(0.6, 0.8)), gravity=2e-4,
                                 method=method, kernel=kernel, iterations=iters, density_ratio=8.0)
        c = c.replace(lattice=int(rng.choice([19, 27])))
    c.body.kernel, c.body.iterations = kernel, iters
    c.body.moving = int(rng.random() < 0.8)
    axes = [0, 2] if two_d else [0, 1, 2]
    periodic = tuple(int(rng.random() < 0.35) if a in axes else 0 for a in range(3))
    c = c.replace(periodic=periodic, name=f"ib_sweep_{seed}")
    R = float(np.abs(c.body.markers).max())
    pos = []
    for a in range(3):
        if a == 1 and two_d:
            pos.append(0.5)
            continue
        lagrange = (stress and a != 0) or (not stress and kernel == _abi.IB_KERNEL_REFERENCE)
        margin = R + (1.2 if (lagrange and not periodic[a]) else 0.1)
        pos.append(float(rng.uniform(margin, dims[a] - margin)))
    c.body.pos = tuple(pos)
    X, Y, Z = c.coords()
    c.init = {"u": 0.01 * np.sin(2 * np.pi * Z / dims[2]), "p": cases.uniform(11 + seed, c.n_cells, -1e-4, 1e-4)}
    return c


@pytest.mark.parametrize("seed", IB_SEEDS)
def test_ib_sweep_against_oracle_fp64(oracle_lib, seed):
    """Fluid, marker forces, nearest cells and the body state after 15 steps, <= 1e-12 (marker
    forces of direct forcing with several passes <= 1e-10: a difference of nearly equal
    velocities)."""
    c = _ib_config(seed)
    with Simulation(c) as sim:
        sim.step(15)
        gu = np.stack([sim.read(x) for x in ("u", "w")])
        gf, gn, gb = sim.markers("force"), sim.markers("node"), sim.body_state()
    o = oracle_lib.Oracle(c)
    o.step(15)
    ou = np.stack([o.read(x) for x in ("u", "w")])
    of, on, ob = o.markers("force"), o.markers("node"), o.body_state()
    o.close()
    b = c.body
    errs = {"vel": rel_l2(gu, ou), "Gw": rel_l2(gf, of.reshape(gf.shape)), "force": rel_l2(gb["force"], ob["force"])}
    if b.moving:
        errs["vel_body"] = rel_l2(gb["vel"], ob["vel"])
    print(c.name, c.lattice, c.n, "periodic", c.periodic, "method", b.method, "kernel", b.kernel, "iters", b.iterations,
          "moving", b.moving, "pos", tuple(round(x, 2) for x in b.pos), errs, flush=True)
    assert np.all(np.isfinite(gu)) and np.all(np.isfinite(gf))
    assert np.array_equal(gn, on)
    for k, e in errs.items():
        assert e <= (1e-10 if (k == "Gw" and b.iterations > 1) else 1e-12), (k, e)


@pytest.mark.parametrize("seed", SEEDS[:16])
def test_sweep_fp32_storage_against_oracle(oracle_lib, seed):
    """The same configurations with FP32 populations (FP32 arithmetic) on both paths: velocity and
    S within 1e-5 of the FP64 oracle after 20 steps."""
    base = _config(seed).replace(storage=F32)
    fields = ["u", "w"] + (["v"] if base.lattice != 9 else []) + (["S"] if base.scalar else [])
    o = oracle_lib.Oracle(base.replace(storage=F64))
    o.step(20)
    want = {f: o.read(f) for f in fields}
    o.close()
    for path in (PATH_PLAIN, PATH_FUSED):
        try:
            sim = Simulation(base.replace(path=path))
        except _abi.IplError:
            assert path == PATH_FUSED
            continue
        with sim:
            sim.step(20)
            got = {f: sim.read(f) for f in fields}
        vel = [f for f in ("u", "v", "w") if f in fields]
        errs = {"vel": rel_l2(np.concatenate([got[f] for f in vel]), np.concatenate([want[f] for f in vel]))}
        if "S" in fields:
            errs["S"] = rel_l2(got["S"], want["S"])
        print(base.name, "fp32 path", path, errs, flush=True)
        for k, e in errs.items():
            assert e <= 1e-5, (path, k, e)


@pytest.mark.parametrize("seed", SEEDS[:24])
def test_sweep_les_and_wall_function_fp64(oracle_lib, seed):
    """The configurations with MRT collision and the LES model (Boegman2005run18IB's coherent-
    structure model, per-cell tau and tau_s) or the wall function's Spalding shear on the cells
    next to walls, drawn per seed (the fused path's EXT step), against the oracle after 20 steps."""
    rng = np.random.default_rng(cases.SEED_BASE + 700 + seed)
    base = _config(seed).replace(collision=_abi.COLL_MRT, path=PATH_FUSED)
    if rng.random() < 0.5:
        mu = float(rng.uniform(0.01, 0.03))
        base = base.replace(les=1, mu_L=mu, mu_H=mu)
        what = "les"
    else:
        base = base.replace(wall_function=1, wall_nu=float(rng.uniform(0.002, 0.01)))
        what = "wall function"
    fields = ["p", "u", "w"] + (["v"] if base.lattice != 9 else []) + (["S"] if base.scalar else [])
    try:
        sim = Simulation(base)
    except _abi.IplError:
        pytest.skip("a configuration the fused path does not take")
    with sim:
        sim.step(20)
        got = {f: sim.read(f) for f in fields}
    o = oracle_lib.Oracle(base)
    o.step(20)
    want = {f: o.read(f) for f in fields}
    o.close()
    vel = [f for f in ("u", "v", "w") if f in fields]
    errs = {"p": rel_l2(got["p"], want["p"]),
            "vel": rel_l2(np.concatenate([got[f] for f in vel]), np.concatenate([want[f] for f in vel]))}
    if "S" in fields:
        errs["S"] = rel_l2(got["S"], want["S"])
    print(base.name, what, errs, flush=True)
    for k, e in errs.items():
        assert e <= 1e-12, (what, k, e)
