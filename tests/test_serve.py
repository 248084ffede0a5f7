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
