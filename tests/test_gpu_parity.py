"""HIP library vs the FP64 oracle on the same seeded inputs.

Tolerance: FP64 storage <= 1e-12 relative L2 after N steps on p, S, phi and on the velocity
vector field (u, v, w) - a single component of a nearly two-dimensional flow (the span-wise v of
the lock exchange) is round-off noise, so velocity is compared as one vector field.  FP32
storage: <= 1e-5 on velocity over these short runs (measured FP32-vs-FP64 drift ~1e-6).
Calls go through the C ABI."""
import numpy as np
import pytest

from ip_cu_lbm_amd import F32, F64, PATH_PLAIN, Simulation, _abi, cases
from conftest import rel_l2

pytestmark = pytest.mark.gpu

FIELDS_FLOW = ("p", "vel")


def _read(x, f):
    if f == "vel":
        return np.stack([x.read(c) for c in "uvw"])
    return x.read(f)


def _compare(case, steps, fields, tol, oracle_lib, report=None):
    o = oracle_lib.Oracle(case)
    with Simulation(case) as sim:
        assert sim.n_calc == o.n_calc and sim.n_ghost == o.n_ghost
        sim.step(steps)
        o.step(steps)
        errs = {f: rel_l2(_read(sim, f), _read(o, f)) for f in fields}
        sums = (sim.global_sums(), o.global_sums())
    print(case.name, steps, errs, flush=True)
    for f, e in errs.items():
        assert e <= tol, (f, e)
    return errs, sums


def test_taylor_green_small_fp64(oracle_lib):
    case = cases.lattice_taylor_green(nx=64, nz=32, path=PATH_PLAIN)
    _compare(case, 300, FIELDS_FLOW, 1e-12, oracle_lib)


def test_taylor_green_full_fp64_1000_steps(oracle_lib):
    case = cases.lattice_taylor_green(path=PATH_PLAIN)
    errs, (sg, so) = _compare(case, 1000, FIELDS_FLOW, 1e-12, oracle_lib)
    assert abs(sg[0] - so[0]) <= 1e-12 * max(1.0, abs(so[0]))


def test_taylor_green_three_point_stencil_fp64(oracle_lib):
    case = cases.lattice_taylor_green(nx=48, nz=16, stencil=3, path=PATH_PLAIN)
    _compare(case, 200, FIELDS_FLOW, 1e-12, oracle_lib)


def test_walls_slip_all_sides_fp64(oracle_lib):
    case = cases.lattice_taylor_green(nx=40, nz=24, path=PATH_PLAIN).replace(periodic=(0, 0, 0), wall_bc=_abi.BC_SLIP)
    _compare(case, 200, FIELDS_FLOW, 1e-12, oracle_lib)


def test_lock_exchange_d3q19_d3q7_fp64(oracle_lib):
    case = cases.lock_exchange(nx=32, ny=6, nz=16, path=PATH_PLAIN)
    _compare(case, 150, ("p", "vel", "S"), 1e-12, oracle_lib)


def test_lock_exchange_same_q_salinity_bounce_back_fp64(oracle_lib):
    case = cases.lock_exchange(nx=24, ny=6, nz=12, scalar_lattice=0, path=PATH_PLAIN).replace(wall_bc=_abi.BC_BOUNCE_BACK)
    _compare(case, 100, ("p", "vel", "S"), 1e-12, oracle_lib)


@pytest.mark.parametrize("q,bc", [(27, _abi.BC_SLIP), (27, _abi.BC_BOUNCE_BACK), (19, _abi.BC_SLIP)])
def test_reference_corner_scheme_3d_fp64(oracle_lib, q, bc):
    nx, ny, nz = 10, 8, 6
    px, wx = cases.axis_from_widths(cases.stretched_widths(nx, 1.1))
    py, wy = cases.axis_from_widths(cases.stretched_widths(ny, 1.15))
    pz, wz = cases.axis_from_widths(np.ones(nz))
    case = cases.Case(name=f"corner_q{q}", lattice=q, n=(nx, ny, nz), pos=[px, py, pz], width=[wx, wy, wz],
                      wall_bc=bc, flux_axes=_abi.FLUX_REFERENCE, tau=0.7, path=PATH_PLAIN,
                      s_rates=np.full(27, 1.2))
    X, Y, Z = case.coords()
    case.init = {"u": 0.01 * np.sin(X / 3.0) * np.cos(Z), "v": 0.01 * np.cos(Y / 2.0),
                 "p": cases.uniform(7, case.n_cells, -1e-3, 1e-3)}
    _compare(case, 60, ("p", "vel"), 1e-12, oracle_lib)


def test_solitary_wave_phase_field_fp64(oracle_lib):
    case = cases.solitary_wave(nx=96, nz=48, x0=30, sigma=1e-4, path=PATH_PLAIN)
    _compare(case, 150, ("p", "vel", "phi"), 1e-12, oracle_lib)


@pytest.mark.parametrize("path", [PATH_PLAIN, _abi.PATH_FUSED])
def test_phase_field_3d_fp64(oracle_lib, path):
    """D3Q19 flow + D3Q19 Allen-Cahn phase field in 3D (surface tension, stretched z, tilted
    interface): the fused path's 3D form takes the scalar-field pass for phi and rho."""
    case = cases.phase_field_3d(n=(24, 8, 18), path=path)
    _compare(case, 100, ("p", "vel", "phi"), 1e-12, oracle_lib)


def test_phase_field_3d_fp32(oracle_lib):
    case = cases.phase_field_3d(n=(24, 8, 18), storage=F32, path=_abi.PATH_FUSED)
    _compare(case, 100, ("vel", "phi"), 1e-5, oracle_lib)


def test_srt_collision_plain_and_fused_fp64(oracle_lib):
    """col_f_SRT (collision.cu:151-165), the collision of the reference's sphere and ellipse drivers
    (SphereFall.cu:356), on both paths: a stretched Taylor-Green channel and the lock exchange."""
    for path in (PATH_PLAIN, _abi.PATH_FUSED):
        tg = cases.lattice_taylor_green(nx=48, nz=24, path=path).replace(collision=_abi.COLL_SRT)
        _compare(tg, 200, FIELDS_FLOW, 1e-12, oracle_lib)
        le = cases.lock_exchange(nx=24, ny=6, nz=12, path=path).replace(collision=_abi.COLL_SRT)
        _compare(le, 100, ("p", "vel", "S"), 1e-12, oracle_lib)


def _sloped_bed(case, h0, slope):
    """Cells below a sloping bed z < h0 + slope x are solid (lnum = -1), the reference's slope
    geometry (Boegman2005run18IB.cu:119-124)."""
    nx, ny, nz = case.n
    x = case.pos[0][:, None, None]
    z = case.pos[2][None, None, :]
    solid = np.broadcast_to(z < h0 + slope * x, (nx, ny, nz)).astype(np.uint8)
    return case.replace(solid=solid.ravel())


@pytest.mark.parametrize("bc", [_abi.BC_BOUNCE_BACK, _abi.BC_SLIP])
def test_cell_mask_sloped_bed_2d_fp64(oracle_lib, bc):
    """A D2Q9 channel stretched in x over a sloping solid bed (ghosts behind every bed cell,
    set_neibghor_wall), plain path, against the oracle."""
    case = cases.lattice_taylor_green(nx=64, nz=32, path=PATH_PLAIN).replace(periodic=(0, 0, 0), wall_bc=bc)
    case = _sloped_bed(case, 2.0, 0.12)
    X, _, Z = case.coords()
    case.init = {"u": 0.01 * np.sin(2 * np.pi * Z / 32.0), "w": 0.005 * np.cos(2 * np.pi * X / 64.0),
                 "p": cases.uniform(11, case.n_cells, -1e-3, 1e-3)}
    assert case.n_cells < 64 * 32
    _compare(case, 200, FIELDS_FLOW, 1e-12, oracle_lib)


def test_cell_mask_lock_exchange_over_a_slope_fp64(oracle_lib):
    """The D3Q19 + D3Q7 lock exchange over a sloping bed (a Boegman-type slope), plain path."""
    case = cases.lock_exchange(nx=32, ny=6, nz=16, path=PATH_PLAIN)
    case = _sloped_bed(case, 1.5, 0.25)
    X, Y, Z = case.coords()
    case.init = {"S": 0.5 * (1.0 - np.tanh((X - 16.0) / 3.0)), "p": cases.uniform(12, case.n_cells, -1e-4, 1e-4)}
    assert case.n_cells < 32 * 6 * 16
    _compare(case, 120, ("p", "vel", "S"), 1e-12, oracle_lib)


def test_taylor_green_fp32_storage(oracle_lib):
    case = cases.lattice_taylor_green(nx=64, nz=32, storage=F32, path=PATH_PLAIN)
    _compare(case, 200, ("vel",), 1e-5, oracle_lib)


def test_nonfinite_detected():
    case = cases.lattice_taylor_green(nx=32, nz=16, path=PATH_PLAIN).replace(check_finite=1)
    with Simulation(case) as sim:
        sim.step(2)
        f = sim.read("f")
        f[3, 17] = np.nan
        sim.write("f", f)
        with pytest.raises(_abi.IplError) as ei:
            sim.step(1)
        assert ei.value.code == _abi.IPL_ERR_NONFINITE


# ---- fused path ---------------------------------------------------------------------------
from ip_cu_lbm_amd import F16, PATH_FUSED  # noqa: E402


def test_fused_taylor_green_fp64(oracle_lib):
    case = cases.lattice_taylor_green(nx=64, nz=32, path=PATH_FUSED)
    _compare(case, 300, FIELDS_FLOW, 1e-12, oracle_lib)


def test_fused_taylor_green_full_fp64(oracle_lib):
    case = cases.lattice_taylor_green(path=PATH_FUSED)
    _compare(case, 1000, FIELDS_FLOW, 1e-12, oracle_lib)


@pytest.mark.parametrize("bc", [_abi.BC_SLIP, _abi.BC_BOUNCE_BACK])
def test_fused_walls_on_stream_axis_fp64(oracle_lib, bc):
    case = cases.lattice_taylor_green(nx=40, nz=24, path=PATH_FUSED).replace(periodic=(0, 0, 0), wall_bc=bc)
    _compare(case, 200, FIELDS_FLOW, 1e-12, oracle_lib)


def test_fused_three_point_fp64(oracle_lib):
    case = cases.lattice_taylor_green(nx=48, nz=16, stencil=3, path=PATH_FUSED)
    _compare(case, 200, FIELDS_FLOW, 1e-12, oracle_lib)


@pytest.mark.parametrize("scalar_lattice,bc", [(7, _abi.BC_SLIP), (0, _abi.BC_BOUNCE_BACK)])
def test_fused_lock_exchange_fp64(oracle_lib, scalar_lattice, bc):
    case = cases.lock_exchange(nx=32, ny=6, nz=16, scalar_lattice=scalar_lattice, path=PATH_FUSED).replace(wall_bc=bc)
    _compare(case, 150, ("p", "vel", "S"), 1e-12, oracle_lib)


def test_fused_uniform_3d_no_flux_fp64(oracle_lib):
    nx, ny, nz = 12, 10, 16
    px, wx = cases.axis_from_widths(np.ones(nx))
    py, wy = cases.axis_from_widths(np.ones(ny))
    pz, wz = cases.axis_from_widths(np.ones(nz))
    case = cases.Case(name="uniform_q19", lattice=19, n=(nx, ny, nz), pos=[px, py, pz], width=[wx, wy, wz],
                      periodic=(1, 0, 0), wall_bc=_abi.BC_BOUNCE_BACK, flux_axes=_abi.FLUX_NONE, tau=0.7,
                      gravity=1e-5, boussinesq=0, path=PATH_FUSED, s_rates=np.full(27, 1.1))
    X, Y, Z = case.coords()
    case.init = {"u": 0.01 * np.sin(2 * np.pi * X / nx), "p": cases.uniform(3, case.n_cells, -1e-3, 1e-3)}
    _compare(case, 100, ("p", "vel"), 1e-12, oracle_lib)


@pytest.mark.parametrize("bc", [_abi.BC_SLIP, _abi.BC_BOUNCE_BACK])
def test_fused_uniform_3d_walls_everywhere_fp64(oracle_lib, bc):
    """Uniform grid (exact streaming), walls on all six faces: the wall redirects of the paired
    pushes at both ends of every line and on the boundary lines."""
    case = cases.lock_exchange(nx=10, ny=6, nz=9, path=PATH_FUSED).replace(periodic=(0, 0, 0), flux_axes=_abi.FLUX_NONE,
                                                                           wall_bc=bc)
    _compare(case, 120, ("p", "vel", "S"), 1e-12, oracle_lib)


def test_fused_solitary_wave_fp64(oracle_lib):
    case = cases.solitary_wave(nx=96, nz=48, x0=30, sigma=1e-4, path=PATH_FUSED)
    _compare(case, 150, ("p", "vel", "phi"), 1e-12, oracle_lib)


def test_fused_lock_exchange_full_size_fp64(oracle_lib):
    """The headline workload at its full size (512 x 32 x 64, z stretched by 1.02, D3Q19 + D3Q7
    salinity) on the fused path against the oracle."""
    case = cases.lock_exchange(path=PATH_FUSED)
    _compare(case, 40, ("p", "vel", "S"), 1e-12, oracle_lib)


def test_fused_solitary_wave_full_size_fp64(oracle_lib):
    """The full 1024 x 256 wave (stretched z, seeded interface noise) over 100 steps.  Longer
    horizons cannot hold 1e-12 for any two FP64 implementations that sum in different orders: the
    oracle against itself with one phi value moved by one ulp differs by 1e-12 in phi after 300
    steps and 4e-4 after 1000 (bench.py --precision --precision-oracle)."""
    case = cases.solitary_wave(path=PATH_FUSED)
    _compare(case, 100, ("p", "vel", "phi"), 1e-12, oracle_lib)


@pytest.mark.parametrize("walls", [False, True])
def test_fused_ragged_sizes_fp64(oracle_lib, walls):
    """Cell counts that are not multiples of 4 or 64 (scalar-field pass tail, partial wavefront
    lines), with and without walls on every face."""
    case = cases.lock_exchange(nx=11, ny=3, nz=9, path=PATH_FUSED)
    if walls:
        case = case.replace(periodic=(0, 0, 0))
    _compare(case, 120, ("p", "vel", "S"), 1e-12, oracle_lib)
    pf = cases.solitary_wave(nx=33, nz=21, x0=10, sigma=1e-4, path=PATH_FUSED)
    _compare(pf, 120, ("p", "vel", "phi"), 1e-12, oracle_lib)


def test_fused_lock_exchange_fp32(oracle_lib):
    case = cases.lock_exchange(nx=32, ny=6, nz=16, storage=F32, path=PATH_FUSED)
    _compare(case, 150, ("vel", "S"), 1e-5, oracle_lib)


def test_fused_taylor_green_fp32(oracle_lib):
    case = cases.lattice_taylor_green(nx=64, nz=32, storage=F32, path=PATH_FUSED)
    _compare(case, 200, ("vel",), 1e-5, oracle_lib)


def test_fused_matches_plain_fp64_lock_exchange():
    base = cases.lock_exchange(nx=32, ny=6, nz=16)
    with Simulation(base.replace(path=PATH_PLAIN)) as a, Simulation(base.replace(path=PATH_FUSED)) as b:
        a.step(100)
        b.step(100)
        assert rel_l2(_read(b, "vel"), _read(a, "vel")) < 1e-12
        assert rel_l2(b.read("S"), a.read("S")) < 1e-12


@pytest.mark.skip(reason="FP16 storage is switched off: its first GPU run faulted and the cause is not yet shown")
def test_fused_lock_exchange_fp16(oracle_lib):
    case = cases.lock_exchange(nx=32, ny=6, nz=16, storage=F16, path=PATH_FUSED)
    _compare(case, 150, ("vel", "S"), 2e-2, oracle_lib)


def _refined_x(coarse=4.0):
    """x refined as the falling-sphere driver does (SphereFall.cu:80-128), in lattice units: coarse
    columns of 4 cells split in 4 near the middle and in 2 around it (finest width = dz)."""
    divs = cases.sphere_fall_divs(10, coarse, inner=(12.0, 28.0), outer=(6.0, 34.0))
    return cases.subdivided_axis(divs, coarse)


def test_fused_refined_axis_fp64(oracle_lib):
    px, wx = _refined_x()
    pz, wz = cases.axis_from_widths(np.ones(16))
    case = cases.Case(name="refined_x", lattice=9, n=(len(px), 1, 16), pos=[px, np.array([0.5]), pz],
                      width=[wx, np.array([1.0]), wz], periodic=(1, 0, 0), wall_bc=_abi.BC_BOUNCE_BACK,
                      flux_axes=_abi.FLUX_X, tau=0.6, path=PATH_FUSED, s_rates=np.array([0, 1.4, 1.4, 0, 1.2, 0, 1.2, 0, 0]))
    X, _, Z = case.coords()
    L = wx.sum()
    case.init = {"u": 0.01 * np.sin(2 * np.pi * X / L) * np.cos(np.pi * Z / 16), "p": cases.uniform(11, case.n_cells, -1e-3, 1e-3)}
    _compare(case, 150, FIELDS_FLOW, 1e-12, oracle_lib)


def test_plain_refined_xy_corner_scheme_fp64(oracle_lib):
    """x and y both refined (divx, divy): the reference's two-axis corner scheme, plain path."""
    px, wx = cases.subdivided_axis([1, 2, 2, 1], 2.0)
    py, wy = cases.subdivided_axis([1, 2, 1], 2.0)
    pz, wz = cases.axis_from_widths(np.ones(5))
    case = cases.Case(name="refined_xy", lattice=27, n=(len(px), len(py), 5), pos=[px, py, pz], width=[wx, wy, wz],
                      wall_bc=_abi.BC_SLIP, flux_axes=_abi.FLUX_REFERENCE, tau=0.7, path=PATH_PLAIN,
                      s_rates=np.full(27, 1.2))
    X, Y, Z = case.coords()
    case.init = {"u": 0.01 * np.sin(X / 2.0), "v": 0.01 * np.cos(Y / 2.0) * np.sin(Z),
                 "p": cases.uniform(12, case.n_cells, -1e-3, 1e-3)}
    _compare(case, 60, ("p", "vel"), 1e-12, oracle_lib)


@pytest.mark.parametrize("bc", [_abi.BC_SLIP, _abi.BC_BOUNCE_BACK])
def test_plain_corner_scheme_bit_identical_runs(bc):
    """The corner scheme's scattered contributions (set_out num3, set_fin) are added with atomics,
    but every destination receives at most two of them in one kernel (its own a0 and the a1 of its
    +x neighbour, or one a2, or its ghost's transfer after set_out has ended), and a sum of two
    terms does not depend on their order: two runs are bit-identical."""
    px, wx = cases.axis_from_widths(cases.stretched_widths(40, 1.03))
    py, wy = cases.axis_from_widths(cases.stretched_widths(36, 1.04))
    pz, wz = cases.axis_from_widths(np.ones(12))
    case = cases.Case(name="corner_repeat", lattice=27, n=(40, 36, 12), pos=[px, py, pz], width=[wx, wy, wz],
                      wall_bc=bc, flux_axes=_abi.FLUX_REFERENCE, tau=0.7, path=PATH_PLAIN, s_rates=np.full(27, 1.2))
    X, Y, Z = case.coords()
    case.init = {"u": 0.01 * np.sin(X / 3.0) * np.cos(Z), "v": 0.01 * np.cos(Y / 2.0),
                 "p": cases.uniform(13, case.n_cells, -1e-3, 1e-3)}
    runs = []
    for _ in range(2):
        with Simulation(case) as sim:
            sim.step(50)
            runs.append(sim.read("f"))
    assert np.all(np.isfinite(runs[0]))
    assert np.array_equal(runs[0], runs[1])


def test_fused_longest_line_fp32(oracle_lib):
    """Lines of 512 cells, the longest the fused path takes (LDS rows of 28 x 512 x 4 B):
    FP32 within 1e-5 of the FP64 oracle; one cell more and AUTO picks the plain path."""
    case = cases.lock_exchange(nx=4, ny=3, nz=512, storage=F32, path=_abi.PATH_AUTO)
    with Simulation(case) as sim:
        assert sim.path == PATH_FUSED
    _compare(case, 40, ("vel", "S"), 1e-5, oracle_lib)
    longer = cases.lock_exchange(nx=2, ny=2, nz=513, storage=F32, path=_abi.PATH_AUTO)
    with Simulation(longer) as sim:
        assert sim.path == PATH_PLAIN
    with pytest.raises(_abi.IplError):
        Simulation(longer.replace(path=PATH_FUSED))


@pytest.mark.parametrize("n,periodic", [((20, 40, 24), (0, 0, 0)), ((40, 24, 20), (0, 1, 0)), ((24, 20, 40), (1, 0, 0))],
                         ids=["lines_along_y", "lines_along_x", "lines_along_z"])
def test_fused_uniform_stream_axis_choice_fp64(oracle_lib, n, periodic):
    """A uniform grid streams along the axis whose lines fill their wavefronts best (40 of 64
    lanes here on y, x and z in turn): every choice matches the oracle.  Walls in z keep p
    hydrostatic: with z periodic it stays near zero and its relative error is ill-conditioned."""
    case = cases.lock_exchange(nx=n[0], ny=n[1], nz=n[2], path=PATH_FUSED).replace(periodic=periodic, flux_axes=_abi.FLUX_NONE)
    _compare(case, 60, ("p", "vel", "S"), 1e-12, oracle_lib)


def test_fused_uniform_stream_axis_lds_budget_fp64(oracle_lib):
    """FP64 lines of 300 cells with a D3Q7 scalar need (19 + 7 + 2) * 300 * 8 = 67,200 bytes of
    LDS, over the 64 KiB budget: the best-filled axis (z) is passed over and the grid streams
    along x, still on the fused path."""
    case = cases.lock_exchange(nx=20, ny=12, nz=300, path=PATH_FUSED).replace(periodic=(0, 1, 0), flux_axes=_abi.FLUX_NONE)
    with Simulation(case) as sim:
        assert sim.path == PATH_FUSED
    _compare(case, 40, ("p", "vel", "S"), 1e-12, oracle_lib)


@pytest.mark.parametrize("lattice,scalar_lattice", [(9, 0), (19, 7), (19, 0)], ids=["d2q9", "d3q7", "d3q19"])
def test_stage_equ_g_fp64(oracle_lib, lattice, scalar_lattice):
    """equ_g as a stage of the plain path (scalar populations <- the equilibrium of S and u)
    against the oracle's, on the same written velocity field; the fused path has no stages."""
    from test_oracle import _random_velocity, _scalar_case
    case = _scalar_case(lattice, scalar_lattice).replace(path=PATH_PLAIN)
    o = oracle_lib.Oracle(case)
    with Simulation(case) as sim:
        vel = _random_velocity(o, o.n_calc, 5)
        for a, x in vel.items():
            sim.write(a, x)
        sim.stage("equ_g")
        o.stage("equ_g")
        got, want = sim.read("g"), o.read("g")
    o.close()
    assert rel_l2(got, want) < 1e-14
    with Simulation(_scalar_case(lattice, scalar_lattice).replace(path=PATH_FUSED)) as fused:
        with pytest.raises(Exception):
            fused.stage("equ_g")
