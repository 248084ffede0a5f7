"""Serving-mix generation and admission sizes on the host (no GPU)."""
import numpy as np

from ip_cu_lbm_amd import _abi, cases, estimate_device_bytes, serve


def test_serving_mix_composition():
    jobs = serve.serving_mix(n_jobs=256)
    assert len(jobs) == 256
    arr = np.array([j.arrival for j in jobs])
    assert np.all(np.diff(arr) > 0) and 3.0 < arr[-1] < 7.5  # mean gap 20 ms
    assert all(500 <= j.steps <= 5000 for j in jobs)
    two_d = [j for j in jobs if j.case.lattice == 9]
    three_d = [j for j in jobs if j.case.lattice == 19]
    assert len(two_d) + len(three_d) == 256 and 90 < len(two_d) < 166
    assert all(64 <= j.case.n[a] <= 512 for j in two_d for a in (0, 2))
    assert all(32 <= j.case.n[a] <= 128 for j in three_d for a in range(3))
    sal = sum(j.case.scalar == _abi.SCALAR_SALINITY for j in jobs)
    ib = [j for j in jobs if j.case.body is not None]
    assert 50 < sal < 110 and 12 < len(ib) < 45
    assert all(200 <= len(j.case.body.markers) <= 1800 for j in ib)
    assert all(j.case.storage == _abi.F32 for j in jobs)
    # seeded: the same stream twice
    again = serve.serving_mix(n_jobs=256)
    assert [(j.case.name, j.steps, j.arrival) for j in jobs] == [(j.case.name, j.steps, j.arrival) for j in again]


def test_cylinder_body():
    c = cases.cylinder_2d(n=(100, 80), markers=300)
    b = c.body
    r = np.linalg.norm(b.markers, axis=1)
    assert np.allclose(r, r[0]) and np.allclose(b.markers[:, 1], 0.0)
    assert np.allclose(b.normals * r[0], b.markers)
    assert abs(b.spacing * 300 - 2 * np.pi * r[0]) < 1e-12
    assert abs(b.mass - b.density * np.pi * r[0] ** 2) < 1e-9


def test_estimate_device_bytes_fused_formula():
    """Fused path, FP32 populations: 2 (Q + Qs) s N for the population pairs plus the force,
    velocity-history and scalar-gradient fields; the estimate grows linearly with cells."""
    a = estimate_device_bytes(cases.lock_exchange(nx=64, ny=16, nz=32, storage=_abi.F32, path=_abi.PATH_FUSED))
    b = estimate_device_bytes(cases.lock_exchange(nx=128, ny=16, nz=32, storage=_abi.F32, path=_abi.PATH_FUSED))
    n = 64 * 16 * 32
    per_cell = (b - a) / n
    assert 2 * (19 + 7) * 4 <= per_cell <= 2 * (19 + 7) * 4 + 160, per_cell
    big = estimate_device_bytes(cases.falling_sphere())
    assert 3.0e9 < big < 4.0e9


def test_server_admission_and_release_accounting(monkeypatch):
    """The serving loop with stand-ins for the library (no GPU): jobs are admitted in arrival order
    under the budget, every admitted job is set up once, stepped only while open and released once
    (on the release thread), and the bytes alive never exceed the budget."""
    import threading
    import time
    import types

    lock = threading.Lock()
    state = {"live": 0, "peak": 0, "created": 0, "closed": 0}
    per_job = 1000

    class FakeSim:
        def __init__(self, case, library=None):
            time.sleep(0.002)
            with lock:
                state["live"] += per_job
                state["peak"] = max(state["peak"], state["live"])
                state["created"] += 1
            self.open = True
            self.steps = 0
            self.prepared = 0

        def prepare(self):
            assert self.open
            self.prepared += 1

        def set_shared(self, on):
            assert self.open and on and not self.prepared  # before the graphs are captured

        def graph_captures(self):
            return 2 * self.prepared

        def close(self):
            time.sleep(0.003)
            with lock:
                assert self.open
                self.open = False
                state["live"] -= per_job
                state["closed"] += 1

    class FakeBatch:
        def __init__(self, library=None, batched=True, max_lines=None):
            self.members = []

        def add(self, sim):
            self.members.append(sim)

        def remove(self, sim):
            self.members.remove(sim)

        def step(self, counts):
            assert len(counts) == len(self.members)
            for sim, k in zip(self.members, counts):
                assert sim.open and k > 0
                sim.steps += k
            time.sleep(0.001)

        def __enter__(self):
            return self

        def __exit__(self, *exc):
            return False

    monkeypatch.setattr(serve, "Simulation", FakeSim)
    monkeypatch.setattr(serve, "Batch", FakeBatch)
    monkeypatch.setattr(serve, "estimate_device_bytes", lambda case, library=None: per_job)
    rng = np.random.default_rng(7)
    jobs, t = [], 0.0
    for j in range(40):
        t += float(rng.exponential(0.002))
        case = types.SimpleNamespace(n_cells=100 + j, name=f"fake{j}")
        jobs.append(serve.Job(jid=j, case=case, steps=int(rng.integers(1, 1200)), arrival=t))
    res = serve.Server(budget_bytes=3.5 * per_job, chunk=400, workers=2).run(jobs)
    assert res["jobs"] == 40
    assert state["created"] == state["closed"] == 40 and state["live"] == 0
    assert state["peak"] <= 3 * per_job
    assert all(j.done_steps == j.steps and j.sim is None and j.finish >= j.start >= j.arrival - 1e-9 for j in jobs)
    assert res["aggregate_mlups"] > 0 and res["p99_s"] >= res["p50_s"] > 0
    assert res["max_graph_captures_per_job"] == 2


def test_estimate_follows_the_path_ipl_create_takes():
    """Under the automatic path the estimate uses the fused formula only where the fused path
    accepts the configuration: a cell mask or the two-axis corner scheme falls back to the plain
    path, and so does its estimate (a larger one)."""
    base = cases.lattice_taylor_green(nx=64, nz=32)
    solid = np.zeros(base.n, dtype=np.uint8)
    solid[10:14, :, :4] = 1
    masked = base.replace(solid=solid.ravel(), init={})
    assert estimate_device_bytes(masked) == estimate_device_bytes(masked.replace(path=_abi.PATH_PLAIN))
    assert estimate_device_bytes(base.replace(init={})) < estimate_device_bytes(masked)
    corner = cases.lock_exchange(nx=12, ny=10, nz=8, scalar_lattice=0).replace(
        lattice=27, scalar=_abi.SCALAR_NONE, flux_axes=_abi.FLUX_REFERENCE, init={})
    assert estimate_device_bytes(corner) == estimate_device_bytes(corner.replace(path=_abi.PATH_PLAIN))


def test_estimate_counts_the_body():
    """A case with an immersed body adds the IB solver's markers, node-block box and body state."""
    c = cases.falling_sphere(n=(40, 40, 60), diameter=10.0, markers=800)
    plain = estimate_device_bytes(c.replace(body=None))
    with_body = estimate_device_bytes(c)
    assert with_body > plain + 800 * 3 * 8 * 6
