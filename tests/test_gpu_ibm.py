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


def test_ib_falling_sphere_full_size_fp64_vs_oracle(oracle_lib):
    """The falling-sphere workload at its full size (192 x 192 x 384 D3Q19, 1800 Fibonacci markers,
    diameter 24, multi-direct forcing with 3 passes and the 4-point Peskin kernel), FP64 on the
    fused path against the oracle after 4 steps (about 30 s on the host's cores): fluid fields,
    marker forces, body force and velocity."""
    case = cases.falling_sphere(storage=F64, path=PATH_FUSED)
    with Simulation(case) as sim:
        sim.step(4)
        gv = np.stack([sim.read(x) for x in "uvw"])
        gp = sim.read("p")
        gf, gb = sim.markers("force"), sim.body_state()
    print("falling sphere full size: GPU done, oracle ...", flush=True)
    o = oracle_lib.Oracle(case)
    try:
        o.step(4)
        ov = np.stack([o.read(x) for x in "uvw"])
        op = o.read("p")
        of, ob = o.markers("force"), o.body_state()
    finally:
        o.close()
    errs = {"vel": rel_l2(gv, ov), "p": rel_l2(gp, op), "Gw": rel_l2(gf, of.reshape(gf.shape)),
            "force": rel_l2(gb["force"], ob["force"]), "vel_body": rel_l2(gb["vel"], ob["vel"])}
    print("falling sphere full size vs oracle, 4 steps:", errs, flush=True)
    for k, e in errs.items():
        assert e <= (1e-10 if k == "Gw" else 1e-12), (k, e)


def test_ib_periodic_axis_fp64(oracle_lib):
    _compare(_sphere(periodic=(1, 0, 0)), 40, oracle_lib, 1e-12)


@pytest.mark.parametrize("method,x", [("stress", 0.6), ("mdf", 0.6), ("mdf", 4.6), ("mdf", 18.9)])
def test_ib_periodic_seam_fp64(oracle_lib, method, x):
    """The body's node blocks across the seam of a periodic axis (below it, x = 0.6 and 4.6, and
    above it, 18.9): the box of the blocks starts at its periodic image, and node blocks, their
    overlaps and the near-marker lists take the nearest periodic image.  At 0.6 and 18.9 some
    markers lie beyond the axis's ends; the nearest-cell search (search_IBlattice, through the
    neighbour table) does not move their node across the seam, so they keep a far node and no
    weight, in the oracle as here."""
    if method == "stress":
        c = _sphere(periodic=(1, 0, 0))
    else:
        c = _sphere(method=_abi.IB_DIRECT, kernel=_abi.IB_KERNEL_PESKIN4, iters=3, density_ratio=8.0, periodic=(1, 0, 0))
    c.body.pos = (x, c.body.pos[1], c.body.pos[2])
    _compare(c, 30, oracle_lib, 1e-12, gw_tol=1e-10)


@pytest.mark.parametrize("method", ["stress", "mdf"])
def test_ib_body_at_a_wall_fp64(oracle_lib, method):
    """Node blocks that reach ghost nodes beyond a wall, which take no weight (and, for multi-
    direct forcing, no axis weight): multi-direct forcing with the body 0.6 cells from the x wall
    and 0.7 from the bottom wall; the stress method at the x wall only, where its weights are the
    masked Peskin function.  (Its y and z weights are 3-point Lagrange over node positions, and a
    ghost node sits on its owner, get_IBMGw2 ibm.cu:146-160: beyond a wall on those axes two of
    the three positions coincide and the weights are not finite, in the oracle as in the
    reference; the drivers keep the body clear of the walls.)"""
    if method == "stress":
        c = _sphere()
        c.body.pos = (4.6, c.body.pos[1], c.body.pos[2])  # radius 4: the markers reach x = 0.6
    else:
        c = _sphere(method=_abi.IB_DIRECT, kernel=_abi.IB_KERNEL_PESKIN4, iters=3, density_ratio=8.0)
        c.body.pos = (4.6, c.body.pos[1], 4.7)
    _compare(c, 20, oracle_lib, 1e-12, gw_tol=1e-10)


def test_ib_short_periodic_axis_mdf_fp64(oracle_lib):
    """A periodic axis of 8 cells, shorter than two node blocks: one block can meet another through
    two periodic images, so the passes stay force + gather pairs."""
    n = (8, 18, 24)
    c = cases.falling_sphere(n=n, diameter=5.0, marker_set="2", storage=F64, path=PATH_FUSED, tau=0.7, gravity=2e-4,
                             method=_abi.IB_DIRECT, kernel=_abi.IB_KERNEL_PESKIN4, iterations=3, density_ratio=8.0)
    c = c.replace(periodic=(1, 0, 0))
    c.body.pos = (4.2, 9.1, 12.3)
    _compare(c, 30, oracle_lib, 1e-12, gw_tol=1e-10)


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


@pytest.mark.parametrize("method", ["stress", "mdf"])
def test_ib_most_markers_fp64(oracle_lib, method):
    """4096 markers, the most one body takes (the rank kernel's LDS), against the oracle; 4097 is
    refused with an error.  With multi-direct forcing every marker has more near markers than its
    list keeps (256), so the marker-space passes scan all markers themselves."""
    n = (28, 28, 30)
    kw = dict(method=_abi.IB_STRESS, kernel=_abi.IB_KERNEL_REFERENCE) if method == "stress" else \
        dict(method=_abi.IB_DIRECT, kernel=_abi.IB_KERNEL_PESKIN4, iterations=3, density_ratio=8.0)
    c = cases.falling_sphere(n=n, diameter=14, markers=4096, storage=F64, path=PATH_FUSED, tau=0.7, gravity=1e-4, **kw)
    c.body.pos = (14.2, 13.9, 15.3)
    _compare(c, 12, oracle_lib, 1e-12, gw_tol=1e-10)
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


def _ib_stagewise(case, oracle_lib, tol, library=None):
    """ipl_ib_stage against the oracle's stages: after a few whole steps, the flow part of a step,
    then each IB kernel in the drivers' order (SphereFall.cu:372-383), comparing what that kernel
    writes: the populations after FLOW, SPM and UPDATE_VEL, marker forces after FORCE, the body
    after BODY, marker positions and nearest cells after POINTS; then whole steps again."""
    stages = ["flow"] + (["spm"] if case.body.spm else []) + ["force", "update_vel", "body", "points"]
    with Simulation(case, library) as sim:
        o = oracle_lib.Oracle(case)
        try:
            sim.step(5)
            o.step(5)
            for rnd in range(2):
                for st in stages:
                    sim.ib_stage(st)
                    o.ib_stage(st)
                    errs = {}
                    if st in ("flow", "spm", "update_vel"):  # these write the populations
                        errs["f"] = rel_l2(sim.read("f"), o.read("f"))
                    if st == "force":
                        errs["Gw"] = rel_l2(sim.markers("force"), o.markers("force"))
                    if st == "body":
                        gb, ob = sim.body_state(), o.body_state()
                        errs["body_force"] = rel_l2(gb["force"], ob["force"])
                        errs["body_vel"] = rel_l2(gb["vel"], ob["vel"])
                    if st == "points":
                        errs["posw"] = rel_l2(sim.markers("pos"), o.markers("pos"))
                        assert np.array_equal(sim.markers("node"), o.markers("node"))
                    print(case.name, rnd, st, errs, flush=True)
                    for k, e in errs.items():
                        assert e <= tol, (rnd, st, k, e)
            sim.step(3)
            o.step(3)
            assert rel_l2(np.stack([sim.read(x) for x in "uvw"]), np.stack([o.read(x) for x in "uvw"])) <= tol
        finally:
            o.close()


def test_ib_stages_stress_method_with_smoothed_profile_fp64(oracle_lib):
    c = _sphere()
    c.body.spm, c.body.spm_a, c.body.spm_b = 1, 4.0, 4.0
    _ib_stagewise(c, oracle_lib, 1e-12)


def test_ib_stages_multi_direct_forcing_fp64(oracle_lib):
    _ib_stagewise(_sphere(method=_abi.IB_DIRECT, kernel=_abi.IB_KERNEL_PESKIN4, iters=3, density_ratio=8.0), oracle_lib, 1e-10)


def test_ib_stages_cylinder_2d_fp64(oracle_lib):
    c = cases.cylinder_2d(n=(48, 56), markers=140, storage=F64, path=PATH_FUSED, density_ratio=3.0, gravity=2e-4)
    _ib_stagewise(c, oracle_lib, 1e-12)


@pytest.mark.parametrize("method", [_abi.IB_STRESS, _abi.IB_DIRECT])
def test_ib_lattice_speed_not_one_fp64(oracle_lib, method):
    """dt = dz / 2, so the lattice speed c = dz / dt = 2: the reference's lattice velocities are
    c e (Items::setc, basic.cu:79-96) in the stress method's split and traction, in
    update_velIBM and in the velocity it leaves (ibm.cu:199-227, 352-358).  Stress method, and
    multi-direct forcing with the Peskin kernel."""
    if method == _abi.IB_STRESS:
        c = _sphere()
    else:
        c = _sphere(method=_abi.IB_DIRECT, kernel=_abi.IB_KERNEL_PESKIN4, iters=3, density_ratio=8.0)
    c = c.replace(dt=0.5, name=c.name + "_c2")
    c.init = {"u": 0.02 * np.sin(2 * np.pi * c.coords()[2] / c.n[2]), "p": cases.uniform(5, c.n_cells, -1e-4, 1e-4)}
    _compare(c, 30, oracle_lib, 1e-12 if method == _abi.IB_STRESS else 1e-10)


def test_ib_stages_lattice_speed_not_one_fp64(oracle_lib):
    c = cases.cylinder_2d(n=(48, 56), markers=140, storage=F64, path=PATH_FUSED, density_ratio=3.0,
                          gravity=2e-4).replace(dt=0.5)
    _ib_stagewise(c, oracle_lib, 1e-12)


@pytest.mark.parametrize("form", [2, 3])
def test_ib_library_smoothed_profile_forms_fp64(oracle_lib, form):
    """The library's SPM_ellipse (form 2, ibm.cu:89-122) and SPM (form 3, ibm.cu:68-87) on a moving,
    spinning sphere (D3Q19, c = 2) and, for SPM_ellipse, a 2D cylinder turned about y."""
    c = _sphere().replace(dt=0.5)
    c.body.omega = (0.0, 0.0, 0.002)
    c.body.spm, c.body.spm_a, c.body.spm_b = form, 3.0, 2.0
    c = c.replace(name=f"{c.name}_spm{form}")
    _compare(c, 25, oracle_lib, 1e-12)
    if form == 2:
        cy = cases.cylinder_2d(n=(48, 56), markers=140, storage=F64, path=PATH_FUSED, density_ratio=3.0, gravity=2e-4)
        cy.body.quat = (np.cos(0.3), 0.0, np.sin(0.3), 0.0)
        cy.body.spm, cy.body.spm_a, cy.body.spm_b = 2, 5.0, 3.0
        _compare(cy.replace(name="cylinder_spm2"), 25, oracle_lib, 1e-12)
