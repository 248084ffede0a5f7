"""Set-up helpers and the open boundary on the host: set_walln, hydrostatic_pressure and the
oracle's set_velocity1 / set_wall_boundary1 restatement (no GPU)."""
import numpy as np
import pytest

from ip_cu_lbm_amd import F64, PATH_FUSED, _abi, cases, hydrostatic_pressure, wall_cells


def _brute_faces(case):
    """Cells next to each non-periodic face, from the cell coordinates (canonical order)."""
    nx, ny, nz = case.n
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ix, iy, iz = ix.ravel(), iy.ravel(), iz.ravel()
    per = case.periodic
    faces = {1: (ix == nx - 1) & (not per[0]), 3: (ix == 0) & (not per[0]),
             5: (iz == nz - 1) & (not per[2]), 6: (iz == 0) & (not per[2])}
    if case.lattice == 9:
        faces[2] = faces[4] = np.zeros(ix.size, bool)
    else:
        faces[2] = (iy == ny - 1) & (not per[1])
        faces[4] = (iy == 0) & (not per[1])
    return {f: np.flatnonzero(m) for f, m in faces.items()}


@pytest.mark.parametrize("make", [
    lambda: cases.lattice_taylor_green(nx=12, nz=8).replace(periodic=(0, 0, 0)),
    lambda: cases.lattice_taylor_green(nx=12, nz=8),
    lambda: cases.lock_exchange(nx=8, ny=5, nz=6),
    lambda: cases.lock_exchange(nx=8, ny=5, nz=6).replace(periodic=(1, 0, 0)),
])
def test_wall_cells_match_faces_and_oracle(oracle_lib, make):
    case = make()
    brute = _brute_faces(case)
    o = oracle_lib.Oracle(case)
    for face in range(1, 7):
        got = wall_cells(case, face)
        assert np.array_equal(got, brute[face]), face
        assert np.array_equal(got, o.wall_cells(face)), face
    o.close()


def test_wall_cells_with_solid_cells(oracle_lib):
    """A solid block inside the box: cells facing it get a ghost too (set_neibghor_wall)."""
    case = cases.lock_exchange(nx=8, ny=4, nz=6, path=_abi.PATH_PLAIN)
    solid = np.zeros(case.n, dtype=np.uint8)
    solid[3:5, :, 2:4] = 1
    case = case.replace(solid=solid.ravel(), init={})
    o = oracle_lib.Oracle(case)
    for face in range(1, 7):
        assert np.array_equal(wall_cells(case, face), o.wall_cells(face)), face
    assert wall_cells(case, 1).size > 4 * 4  # the +x wall plus the block's -x face
    o.close()


@pytest.mark.parametrize("boussinesq", [True, False])
def test_hydrostatic_pressure_closed_form(oracle_lib, boussinesq):
    """p = 3 g (nz dz / 2 - z) / c^2 / 1000 (Boussinesq) or 3 g (z0 - z) / c^2, g = 9.81, on a
    z-stretched grid; the library and the oracle's restatement agree."""
    case = cases.solitary_wave(nx=16, nz=24, x0=5)
    _, _, Z = case.coords()
    z = Z.ravel()
    c = case.dz / case.dt
    want = (3 * 9.81 * (24 / 2.0 * case.dz - z) / c ** 2 / 1000) if boussinesq else 3 * 9.81 * (7.5 - z) / c ** 2
    got = hydrostatic_pressure(case, boussinesq, 7.5)
    assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()
    o = oracle_lib.Oracle(case)
    assert np.abs(o.hydrostatic_pressure(boussinesq, 7.5) - got).max() <= 1e-13 * np.abs(want).max()
    o.close()


def test_open_boundary_keeps_uniform_stream(oracle_lib):
    """A uniform stream U through a channel with slip walls, inlet and outlet columns held at
    feq(0, U): every cell keeps u = U, w = 0 (the open boundary feeds the stream it removes)."""
    U = 0.02
    case = cases.lattice_taylor_green(nx=24, nz=10, path=PATH_FUSED).replace(periodic=(0, 0, 0), wall_bc=_abi.BC_SLIP)
    n = case.n_cells
    case.init = {"u": np.full(n, U), "p": np.zeros(n)}
    o = oracle_lib.Oracle(case)
    cells = np.concatenate([wall_cells(case, 3), wall_cells(case, 1)])
    o.set_open_boundary(cells, 0.0, U, 0.0, 0.0)
    o.step(60)
    assert np.abs(o.read("u") - U).max() < 1e-13
    assert np.abs(o.read("w")).max() < 1e-13
    # and it is the equilibrium the reference writes (set_wall_boundary1): f = w (p + 3t + 4.5t^2 - 1.5u^2)
    f = o.read("f").reshape(9, n)
    L = [(0, 0), (1, 0), (0, 1), (-1, 0), (0, -1), (1, 1), (-1, 1), (-1, -1), (1, -1)]
    W = [4 / 9] + [1 / 9] * 4 + [1 / 36] * 4
    for k, ((ex, ez), w) in enumerate(zip(L, W)):
        t = ex * U
        assert np.abs(f[k, cells] - w * (3 * t + 4.5 * t * t - 1.5 * U * U)).max() < 1e-15
    o.close()


@pytest.mark.parametrize("abs_u,y,nu", [(0.01, 0.5, 1e-3), (0.05, 0.5, 1e-4), (0.002, 1.0, 1e-5), (0.1, 0.25, 5e-5)])
def test_wall_function_newton_solves_spalding_law(oracle_lib, abs_u, y, nu):
    """The wall function's u_tau (15 Newton steps, collision.cu:37-50) is the root of Spalding's
    law y+ = u+ + e^{-kB}(e^{k u+} - 1 - k u+ - (k u+)^2/2 - (k u+)^3/6), k = 0.41, B = 5.5, as
    scipy's bracketing solver finds it, within the offset of the reference's regularisation
    (eta = 1e-7 |u| added to u_tau in u+ and to the derivative): 2e-6 relative, for y+ from 2 to
    37 where 15 steps converge (far into the log layer they do not, in the reference as here)."""
    from scipy.optimize import brentq
    k, B = 0.41, 5.5

    def resid(ut):
        up = abs_u / ut
        return y * ut / nu - (up + np.exp(-k * B) * (np.exp(k * up) - 1 - k * up - (k * up) ** 2 / 2 - (k * up) ** 3 / 6))

    ref = brentq(resid, 1e-8 * abs_u, abs_u, xtol=1e-16, rtol=1e-15)
    got = oracle_lib.spalding_utau(abs_u, y, nu)
    assert abs(got - ref) <= 2e-6 * ref, (got, ref)


def test_les_closed_form(oracle_lib):
    """LES (basic.cu:285-339) on u = a x, v = b y, w = c z: the central differences are exact, so
    at every interior cell nabla = diag(a, b, c), Q = -(a^2 + b^2 + c^2)/2, E = 1/(2 (a+b+c)^2),
    F = clip(Q/E), C = 0.05 |F|^(3/2), S = sqrt(2) (|a| + |b| + |c|), tau = 3 (nu + C S dz^2)/(c^2 dt)
    + 1/2 and tau_s = 3 (1e-9 + C S dz^2 / 1000)/(c^2 dt) + 1/2, with nu = mu_L / rho."""
    a, b, c = 0.3, -0.2, 0.5
    case = cases.lock_exchange(nx=8, ny=7, nz=9).replace(flux_axes=_abi.FLUX_NONE, les=1, mu_L=0.01, mu_H=0.01,
                                                        scalar=_abi.SCALAR_NONE, scalar_lattice=0, gravity=0.0)
    pos = [cases.axis_from_widths(np.ones(n))[0] for n in case.n]
    case = case.replace(pos=pos, width=[np.ones(n) for n in case.n])
    X, Y, Z = case.coords()
    case.init = {"u": a * X.ravel(), "v": b * Y.ravel(), "w": c * Z.ravel()}
    o = oracle_lib.Oracle(case)
    o.les()
    tau, taus = o.read("tau").reshape(case.n), o.read("taus").reshape(case.n)
    o.close()
    Q = -(a * a + b * b + c * c) / 2
    E = 1 / (2 * (a + b + c) ** 2 + 1e-40)
    F = max(-1.0, min(Q / E, 1.0))
    C = 0.05 * abs(F) ** 1.5
    S = np.sqrt(2) * (abs(a) + abs(b) + abs(c))
    # the grid's own dz/1000 guard in the differences: (v+ - v-)/(2 + dz/1000)
    g = 2.0 / (2.0 + 1e-3)
    Qg, Eg = Q * g * g, 1 / (2 * ((a + b + c) * g) ** 2 + 1e-40)
    Cg = 0.05 * abs(max(-1.0, min(Qg / Eg, 1.0))) ** 1.5
    Sg = S * g
    want_tau = 3 * (0.01 / 1.0 + Cg * Sg) + 0.5
    want_taus = 3 * (1e-9 + Cg * Sg / 1000) + 0.5
    inner = (slice(1, -1),) * 3
    assert np.abs(tau[inner] - want_tau).max() < 1e-14
    assert np.abs(taus[inner] - want_taus).max() < 1e-15
    assert Cg * Sg > 1e-6  # the eddy part is visible at this size
