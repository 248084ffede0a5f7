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

from ip_cu_lbm_amd import F64, PATH_FUSED, PATH_PLAIN, Simulation, _abi, cases
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
