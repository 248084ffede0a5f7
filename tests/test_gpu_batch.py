"""Graph replay, batched stepping and the serving loop on the GPU.

A step replayed from a captured HIP graph runs the same kernels with the same arguments as a
directly launched one, and members of a batch share nothing but the device, so every comparison
here is bit-exact: graph vs direct, each job alone vs the same job inside a batch or a served mix.
Each graph test replays several times with reads (device-to-host copies and synchronizations)
between the replays and checks the state after every one of them."""
import numpy as np
import pytest

from ip_cu_lbm_amd import F32, F64, PATH_FUSED, PATH_PLAIN, Batch, Simulation, _abi, cases, estimate_device_bytes, serve

pytestmark = pytest.mark.gpu


def _state(sim, fields=("p", "u", "v", "w")):
    return {f: sim.read(f) for f in fields}


def _same(a, b, what):
    for k in a:
        assert np.array_equal(a[k], b[k]), (what, k, np.abs(a[k] - b[k]).max())


def _small_cases():
    le = cases.lock_exchange(nx=40, ny=6, nz=18, storage=F32, path=PATH_FUSED)
    tg = cases.lattice_taylor_green(nx=96, nz=48, storage=F32, path=PATH_FUSED)
    pf = cases.solitary_wave(nx=64, nz=32, x0=20, sigma=1e-4, path=PATH_FUSED)
    cyl = cases.cylinder_2d(n=(72, 80), markers=160, storage=F32, path=PATH_FUSED)
    sph = cases.falling_sphere(n=(24, 22, 30), diameter=8.0, marker_set="3", storage=F32, path=PATH_FUSED,
                               method=_abi.IB_STRESS, kernel=_abi.IB_KERNEL_REFERENCE, tau=0.7, gravity=2e-4)
    sph.body.pos = (12.3, 10.8, 15.4)
    return {"le": (le, ("p", "u", "v", "w", "S")), "tg": (tg, ("p", "u", "w")), "pf": (pf, ("p", "u", "w", "phi")),
            "cyl": (cyl, ("p", "u", "w")), "sph": (sph, ("p", "u", "v", "w"))}


@pytest.mark.parametrize("name", ["le", "tg", "pf", "cyl", "sph"])
def test_graph_replay_matches_direct_launches(name):
    """Graph replay (20-step graphs) vs direct launches, bit for bit, after each of several
    replays; an odd direct count in between swaps the buffer roles, which switches to the other
    parity's graph: each of the two graphs is captured once."""
    case, fields = _small_cases()[name]
    with Simulation(case) as g, Simulation(case) as d:
        d.set_graph(False)
        for n in (20, 20, 40, 7, 20, 60, 3, 21, 41, 20):
            g.step(n)
            d.step(n)
            _same(_state(g, fields), _state(d, fields), (name, n))
        assert g.graph_captures() == 2
        assert d.graph_captures() == 0


def test_prepare_captures_both_graphs_once():
    """ipl_prepare captures the graphs of both buffer parities before any step; later steps of
    any count capture nothing more, and the results equal direct launches bit for bit."""
    case, fields = _small_cases()["le"]
    with Simulation(case) as g, Simulation(case) as d:
        d.set_graph(False)
        g.prepare()
        assert g.graph_captures() == 2
        g.prepare()
        for n in (5, 20, 33, 40):
            g.step(n)
            d.step(n)
        assert g.graph_captures() == 2
        _same(_state(g, fields), _state(d, fields), "prepare")
        if case.body is not None:
            bg, bd = g.body_state(), d.body_state()
            for k in bg:
                assert np.array_equal(bg[k], bd[k]), k
            assert np.array_equal(g.markers("force"), d.markers("force"))


def test_batch_members_match_alone():
    """Every member of a batch, advanced by its own step counts, ends bit-identical to the same
    case stepped alone; members join and leave between batch steps."""
    cs = _small_cases()
    names = list(cs)
    counts = {"le": [30, 25, 40], "tg": [50, 0, 13], "pf": [20, 20, 20], "cyl": [7, 40, 33], "sph": [25, 25, 5]}
    alone = {}
    for n in names:
        case, fields = cs[n]
        with Simulation(case) as s:
            s.step(sum(counts[n]))
            alone[n] = _state(s, fields)
    sims = {n: Simulation(cs[n][0]) for n in names}
    try:
        with Batch() as b:
            for n in names[:3]:
                b.add(sims[n])
            b.step([counts[n][0] for n in names[:3]])
            for n in names[3:]:
                b.add(sims[n])
            # the late joiners catch up with a first chunk of their own
            b.step([counts[n][1] for n in names[:3]] + [counts[n][0] for n in names[3:]])
            b.remove(sims["tg"])
            sims["tg"].step(counts["tg"][2])
            rest = [n for n in names if n != "tg"]
            b.step([counts[n][2] if n in names[:3] else counts[n][1] for n in rest])
            b.step([0 if n in names[:3] else counts[n][2] for n in rest])
        for n in names:
            _same(_state(sims[n], cs[n][1]), alone[n], n)
    finally:
        for s in sims.values():
            s.close()


def test_batch_uniform_count():
    cs = _small_cases()
    with Simulation(cs["le"][0]) as a, Simulation(cs["tg"][0]) as b, Simulation(cs["le"][0]) as ref:
        with Batch() as bt:
            bt.add(a)
            bt.add(b)
            bt.step(45)
        ref.step(45)
        _same(_state(a, cs["le"][1]), _state(ref, cs["le"][1]), "uniform count")


def _launch_mix():
    """Members of several batched-launch groups: D2Q9 FP32 at three sizes whose lines round to one
    workgroup size and one that does not, D3Q19+D3Q7 FP32 at two sizes, D2Q9 FP64, a phase-field
    D2Q9+D2Q9 wave, an immersed-boundary cylinder and a falling sphere with multi-direct forcing
    (stepped on their own streams, the body's step tail on the member's stream: alone it runs on a
    second stream)."""
    f2 = ("p", "u", "w")
    return [(cases.lattice_taylor_green(nx=96, nz=48, storage=F32, path=PATH_FUSED), f2, True),
            (cases.lattice_taylor_green(nx=80, nz=72, storage=F32, path=PATH_FUSED), f2, True),
            (cases.lattice_taylor_green(nx=120, nz=40, storage=F32, path=PATH_FUSED), f2, True),
            (cases.lattice_taylor_green(nx=300, nz=200, storage=F32, path=PATH_FUSED), f2, True),
            (cases.lock_exchange(nx=40, ny=6, nz=18, storage=F32, path=PATH_FUSED), ("p", "u", "v", "w", "S"), True),
            (cases.lock_exchange(nx=36, ny=10, nz=20, storage=F32, path=PATH_FUSED), ("p", "u", "v", "w", "S"), True),
            (cases.lattice_taylor_green(nx=64, nz=40, storage=F64, path=PATH_FUSED), f2, True),
            (cases.solitary_wave(nx=64, nz=32, x0=20, sigma=1e-4, path=PATH_FUSED), ("p", "u", "w", "phi"), True),
            (cases.cylinder_2d(n=(72, 80), markers=160, storage=F32, path=PATH_FUSED), f2, False),
            (cases.falling_sphere(n=(20, 18, 24), diameter=8.0, marker_set="3", storage=F32, path=PATH_FUSED, tau=0.7,
                                  gravity=2e-4, density_ratio=8.0), ("p", "u", "v", "w"), False)]


def test_batched_launches_match_alone():
    """Batched launches (one k_fused_step launch per group and step) against each job
    stepped alone, bit for bit, over rounds of uneven counts with zeros and odd counts (which run
    as phases of decreasing membership and leave the buffer roles swapped); the counters show
    which member-steps the batched launches advanced.  Per-member launches give the same."""
    mix = _launch_mix()
    rounds = [[30, 25, 40, 7, 19, 1, 13, 22, 9, 21], [0, 31, 2, 40, 40, 40, 5, 0, 11, 20],
              [17, 17, 17, 17, 3, 0, 28, 30, 6, 3]]
    alone = []
    for i, (case, fields, _b) in enumerate(mix):
        with Simulation(case) as s:
            s.step(sum(r[i] for r in rounds))
            alone.append(_state(s, fields))
    for batched in (True, False):
        sims = [Simulation(c) for c, _f, _b in mix]
        try:
            with Batch(batched=batched) as b:
                for s in sims:
                    b.add(s)
                for r in rounds:
                    b.step(r)
                ctr = b.counters()
            for i, (case, fields, _b) in enumerate(mix):
                _same(_state(sims[i], fields), alone[i], (case.name, batched))
        finally:
            for s in sims:
                s.close()
        print(batched, ctr, flush=True)
        in_batch = sum(r[i] for r in rounds for i in range(len(mix)) if mix[i][2])
        total = sum(sum(r) for r in rounds)
        if batched:
            assert ctr["batched_steps"] == in_batch and ctr["own_steps"] == total - in_batch, ctr
            assert 0 < ctr["launches"] < ctr["batched_steps"], ctr
        else:
            assert ctr["batched_steps"] == 0 and ctr["launches"] == 0 and ctr["own_steps"] == total, ctr


def test_batched_graphs_reused_per_layout():
    """A repeated pattern of step calls (same members, same counts) reuses the graphs captured on
    its first pass: one per phase layout; results after every call equal direct launches."""
    mix = [m for m in _launch_mix() if m[2]][:5]
    pattern = [[20, 20, 20, 20, 20], [8, 8, 4, 4, 8], [6, 2, 6, 2, 6]]
    sims = [Simulation(c) for c, _f, _b in mix]
    refs = [Simulation(c) for c, _f, _b in mix]
    try:
        with Batch() as b:
            for s in sims:
                b.add(s)
            captures = []
            for rep in range(3):
                for r in pattern:
                    b.step(r)
                    for ref, k in zip(refs, r):
                        ref.step(k)
                    for i, (case, fields, _b) in enumerate(mix):
                        _same(_state(sims[i], fields), _state(refs[i], fields), (case.name, rep, r))
                captures.append(b.counters()["graph_captures"])
            print(captures, b.counters(), flush=True)
            assert captures[0] > 0 and captures[2] == captures[1] == captures[0], captures
    finally:
        for s in sims + refs:
            s.close()


def test_estimate_covers_allocation():
    """The admission estimate is within 25 % above what the context actually allocates."""
    for name, (case, _f) in _small_cases().items():
        est = estimate_device_bytes(case)
        with Simulation(case) as s:
            real = s.device_bytes
        print(name, est, real, flush=True)
        assert real <= est <= 1.25 * real + (2 << 20), (name, est, real)


def test_serving_mix_results_match_alone():
    """A small served mix (continuous batching, memory budget forcing a queue): every job's final
    fields equal the same job run alone."""
    jobs = serve.serving_mix(n_jobs=14, mean_gap_s=0.002, steps=(40, 160), seed=cases.SEED_BASE + 11)
    finals = {}

    def keep(job):
        finals[job.jid] = _state(job.sim, ("p", "u", "w"))

    budget = 3 * max(estimate_device_bytes(j.case) for j in jobs)
    res = serve.Server(budget_bytes=budget, chunk=30).run(jobs, log=print, on_done=keep)
    print(res, flush=True)
    assert res["jobs"] == len(jobs) and len(finals) == len(jobs)
    assert res["p99_s"] >= res["p50_s"] > 0 and res["aggregate_mlups"] > 0
    for j in jobs:
        with Simulation(j.case) as s:
            s.step(j.steps)
            _same(_state(s, ("p", "u", "w")), finals[j.jid], j.case.name)


def test_cylinder_2d_stress_fp64(oracle_lib):
    """2D immersed boundary (D2Q9, markers on a circle in the x-z plane) vs the oracle.  Multi-direct
    forcing with the reference's 3 x 3 block repeats force + gather per pass (its node weights are not
    products of axis weights); with the Peskin block the later passes run in marker space."""
    from conftest import rel_l2
    c = cases.cylinder_2d(n=(48, 56), markers=140, storage=F64, path=PATH_FUSED, density_ratio=3.0, gravity=2e-4)
    c.body.pos = (24.3, 0.5, 27.6)
    X, _, Z = c.coords()
    c.init = {"u": 0.01 * np.sin(2 * np.pi * Z / 56), "p": cases.uniform(9, c.n_cells, -1e-4, 1e-4)}
    for method, kernel, iters in ((_abi.IB_STRESS, _abi.IB_KERNEL_REFERENCE, 1),
                                  (_abi.IB_DIRECT, _abi.IB_KERNEL_REFERENCE, 1),
                                  (_abi.IB_DIRECT, _abi.IB_KERNEL_REFERENCE, 3),
                                  (_abi.IB_DIRECT, _abi.IB_KERNEL_PESKIN4, 3)):
        c.body.method, c.body.kernel, c.body.iterations = method, kernel, iters
        with Simulation(c) as sim:
            sim.step(40)
            gb, gf, gu = sim.body_state(), sim.markers("force"), np.stack([sim.read("u"), sim.read("w")])
        o = oracle_lib.Oracle(c)
        o.step(40)
        ob, of, ou = o.body_state(), o.markers("force"), np.stack([o.read("u"), o.read("w")])
        o.close()
        errs = (rel_l2(gu, ou), rel_l2(gf, of), rel_l2(gb["force"], ob["force"]), rel_l2(gb["vel"], ob["vel"]))
        print("cylinder", method, kernel, iters, errs, flush=True)
        assert errs[0] <= 1e-12 and errs[2] <= 1e-12 and errs[3] <= 1e-12
        assert errs[1] <= (1e-10 if iters > 1 else 1e-12)


@pytest.mark.parametrize("x", [4.9, 43.1])
@pytest.mark.parametrize("method,kernel,iters", [(_abi.IB_STRESS, _abi.IB_KERNEL_REFERENCE, 1),
                                                 (_abi.IB_DIRECT, _abi.IB_KERNEL_PESKIN4, 3)])
def test_cylinder_2d_across_a_periodic_seam_fp64(oracle_lib, method, kernel, iters, x):
    """A 2D body whose node blocks cross the seam of a periodic x axis (radius 4.8: markers from
    x = 0.1, or up to 47.9 of 48): the blocks wrap, the box of node blocks starts at the seam's
    image, and the marker-space passes take the nearest periodic image.  (Markers themselves stay
    inside the axis: the nearest-cell search, as search_IBlattice through the neighbour table,
    does not move a marker's node across the seam, so a marker beyond it would sit far from its
    node, in the oracle as here; the 2D stress method's renormalised weights are then 0/0.)"""
    from conftest import rel_l2
    c = cases.cylinder_2d(n=(48, 56), markers=140, storage=F64, path=PATH_FUSED, density_ratio=3.0, gravity=2e-4)
    c = c.replace(periodic=(1, 0, 0))
    c.body.pos = (x, 0.5, 27.6)
    c.body.method, c.body.kernel, c.body.iterations = method, kernel, iters
    X, _, Z = c.coords()
    c.init = {"u": 0.01 * np.sin(2 * np.pi * Z / 56), "p": cases.uniform(9, c.n_cells, -1e-4, 1e-4)}
    with Simulation(c) as sim:
        sim.step(40)
        gb, gf, gu = sim.body_state(), sim.markers("force"), np.stack([sim.read("u"), sim.read("w")])
    o = oracle_lib.Oracle(c)
    o.step(40)
    ob, of, ou = o.body_state(), o.markers("force"), np.stack([o.read("u"), o.read("w")])
    o.close()
    errs = (rel_l2(gu, ou), rel_l2(gf, of), rel_l2(gb["force"], ob["force"]), rel_l2(gb["vel"], ob["vel"]))
    print("cylinder across the seam", method, kernel, iters, errs, flush=True)
    assert errs[0] <= 1e-12 and errs[2] <= 1e-12 and errs[3] <= 1e-12
    assert errs[1] <= (1e-10 if iters > 1 else 1e-12)


def test_path_reason_names_the_fallback():
    """Under the automatic path a configuration the fused path refuses runs on the plain path, and
    the context says why (ipl_path_reason); a fused context has no reason."""
    base = cases.lattice_taylor_green(nx=32, nz=16)
    solid = np.zeros(base.n, dtype=np.uint8)
    solid[5:8, :, :3] = 1
    with Simulation(base.replace(solid=solid.ravel(), init={})) as s:
        assert s.path == PATH_PLAIN and "mask" in s.path_reason
    with Simulation(base) as s:
        assert s.path == PATH_FUSED and s.path_reason == ""
