"""Immersed boundary on the GPU vs the FP64 oracle (function/ibm.cu, SphereFall.cu:372-383).

Same seeded fluid state and body on both sides; after N steps the fluid fields, the marker
positions and forces Gw, the nearest cells and the body state (force, torque, velocity) agree to
<= 1e-12 relative L2 in FP64.  The spread on the GPU sums each cell's contributions in a fixed
order (slots, then markers by index), so two runs are bit-identical; the oracle sums in marker
order, which differs only by rounding."""
import numpy as np
import pytest

from ip_cu_lbm_amd import F32, F64, PATH_FUSED, Simulation, _abi, cases
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _sphere(method=_abi.IB_STRESS, kernel=_abi.IB_KERNEL_REFERENCE, iters=1, q=19, storage=F64, moving=1,
            n=(20, 18, 24), periodic=(0, 0, 0), ico=3, density_ratio=1.5):
    c = cases.falling_sphere(n=n, diameter=8.0, method=method, kernel=kernel, iterations=iters, storage=storage,
                             path=PATH_FUSED, marker_set=str(ico), gravity=2e-4, tau=0.7, density_ratio=density_ratio)
    c = c.replace(lattice=q, periodic=periodic, name=f"sphere_q{q}_m{method}_k{kernel}_i{iters}")
    c.body.pos = (n[0] / 2 + 0.3, n[1] / 2 - 0.2, n[2] / 2 + 0.4)
    c.body.moving = moving
    X, Y, Z = c.coords()
    c.init = {"u": 0.01 * np.sin(2 * np.pi * Z / n[2]), "p": cases.uniform(5, c.n_cells, -1e-4, 1e-4)}
    return c


def _compare(case, steps, oracle_lib, tol, gw_tol=None):
    with Simulation(case) as sim:
        sim.step(steps)
        gb = sim.body_state()
        gm = {k: sim.markers(k) for k in ("pos", "force", "node")}
        gv = np.stack([sim.read(x) for x in "uvw"])
        gp = sim.read("p")
    o = oracle_lib.Oracle(case)
    o.step(steps)
    ob = o.body_state()
    om = {k: o.markers(k) for k in ("pos", "force", "node")}
    ov = np.stack([o.read(x) for x in "uvw"])
    op = o.read("p")
    o.close()
    errs = {"vel": rel_l2(gv, ov), "p": rel_l2(gp, op), "Gw": rel_l2(gm["force"], om["force"]),
            "posw": rel_l2(gm["pos"], om["pos"]), "force": rel_l2(gb["force"], ob["force"]),
            "vel_body": rel_l2(gb["vel"], ob["vel"])}
    print(case.name, steps, errs, flush=True)
    for k, e in errs.items():
        assert e <= (gw_tol if (k == "Gw" and gw_tol) else tol), (k, e)
    assert np.array_equal(gm["node"], om["node"])
    return gb, ob


@pytest.mark.parametrize("q", [19, 27])
def test_ib_stress_method_fp64(oracle_lib, q):
    _compare(_sphere(q=q), 40, oracle_lib, 1e-12)


def test_ib_stress_method_srt_fp64(oracle_lib):
    """The falling sphere with SRT collision, as the reference's sphere driver runs it
    (SphereFall.cu:356, col_f_SRT collision.cu:151-165)."""
    c = _sphere()
    _compare(c.replace(collision=_abi.COLL_SRT, name=c.name + "_srt"), 40, oracle_lib, 1e-12)


def test_ib_direct_forcing_fp64(oracle_lib):
    _compare(_sphere(method=_abi.IB_DIRECT), 40, oracle_lib, 1e-12)


def test_ib_multi_direct_forcing_peskin4_fp64(oracle_lib):
    """Three forcing passes with the full 4-point Peskin support.  On this small sphere (shell volume
    0.75 of the body volume) the reference's explicit body update is stable only for heavy bodies,
    so the body is 8 times denser than the fluid; the C3-size sphere below runs at 1.5."""
    _compare(_sphere(method=_abi.IB_DIRECT, kernel=_abi.IB_KERNEL_PESKIN4, iters=3, density_ratio=8.0), 40, oracle_lib,
             1e-12, gw_tol=1e-10)


def test_ib_falling_sphere_d16_fp64(oracle_lib):
    """Sphere of diameter 16 on ~800 Fibonacci markers, density ratio 1.5, multi-direct forcing (3
    passes, 4-point Peskin): the falling-sphere workload at reduced size."""
    n = (28, 28, 34)
    c = cases.falling_sphere(n=n, diameter=16, markers=804, storage=F64, path=PATH_FUSED, tau=0.644, gravity=1e-4)
    c.body.pos = (n[0] / 2 + 0.25, n[1] / 2 - 0.1, n[2] / 2 + 2)
    # After three passes the marker force (U - u_m)/dt is a difference of nearly equal velocities,
    # so its relative rounding error is amplified (the summed body force is not): 1e-10 on Gw.
    gb, ob = _compare(c, 60, oracle_lib, 1e-12, gw_tol=1e-10)
    assert gb["vel"][2] < 0  # falling along -z


def test_ib_periodic_axis_fp64(oracle_lib):
    _compare(_sphere(periodic=(1, 0, 0)), 40, oracle_lib, 1e-12)


def test_ib_fixed_body_fp64(oracle_lib):
    gb, ob = _compare(_sphere(moving=0), 40, oracle_lib, 1e-12)
    assert np.allclose(gb["pos"], ob["pos"]) and np.allclose(gb["vel"], 0.0)


def test_ib_fp32_storage(oracle_lib):
    """FP32 populations with FP64 marker and body arithmetic: force and body velocity within 1e-5."""
    _compare(_sphere(storage=F32), 40, oracle_lib, 1e-5)


def test_ib_deterministic():
    """Two runs of the same case give bit-identical body and marker states (no atomics)."""
    case = _sphere(method=_abi.IB_DIRECT, kernel=_abi.IB_KERNEL_PESKIN4, iters=3, density_ratio=8.0)
    out = []
    for _ in range(2):
        with Simulation(case) as sim:
            sim.step(30)
            out.append((sim.body_state()["force"].copy(), sim.markers("force").copy(), sim.read("u")))
    assert np.all(np.isfinite(out[0][0]))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2], out[1][2])


@pytest.mark.parametrize("q", [19, 27])
def test_ib_smoothed_profile_fp64(oracle_lib, q):
    """Stress method with the driver's smoothed-profile interior forcing (SPM_ellipse3D,
    SphereFall3D/SphereFall.cu:13-30, called before get_IBMGw2 at :376): semi-axes R, R as there."""
    c = _sphere(q=q)
    c.body.spm, c.body.spm_a, c.body.spm_b = 1, 4.0, 4.0
    gb, ob = _compare(c, 30, oracle_lib, 1e-12)
    c0 = _sphere(q=q)
    with Simulation(c0) as s:
        s.step(30)
        assert np.abs(s.body_state()["force"] - gb["force"]).max() > 1e-8  # the interior forcing acts


def test_ib_most_markers_fp64(oracle_lib):
    """4096 markers, the most one body takes (the rank kernel's LDS), against the oracle; 4097 is
    refused with an error."""
    n = (28, 28, 30)
    c = cases.falling_sphere(n=n, diameter=14, markers=4096, storage=F64, path=PATH_FUSED, tau=0.7, gravity=1e-4,
                             method=_abi.IB_STRESS, kernel=_abi.IB_KERNEL_REFERENCE)
    c.body.pos = (14.2, 13.9, 15.3)
    _compare(c, 12, oracle_lib, 1e-12)
    over = cases.falling_sphere(n=n, diameter=14, markers=4097, storage=F64, path=PATH_FUSED,
                                method=_abi.IB_STRESS, kernel=_abi.IB_KERNEL_REFERENCE)
    over.body.pos = (14.2, 13.9, 15.3)
    with pytest.raises(_abi.IplError):
        Simulation(over)


def test_ib_lines_along_y_fp64(oracle_lib):
    """Multi-direct forcing on a uniform grid whose lines run along y (the best-filled axis)."""
    c = cases.falling_sphere(n=(20, 40, 24), diameter=8.0, marker_set="3", storage=F64, path=PATH_FUSED, tau=0.7,
                             gravity=2e-4, method=_abi.IB_DIRECT, kernel=_abi.IB_KERNEL_PESKIN4, iterations=3,
                             density_ratio=8.0)
    c.body.pos = (10.3, 20.2, 12.4)
    _compare(c, 30, oracle_lib, 1e-12, gw_tol=1e-10)
