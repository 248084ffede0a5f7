"""The HIP library's host-side graph builder against the oracle's restatement of the
reference builder (neighbour table with ghosts, slip remap tables, stencil weights).
Host-only: the library is loaded but no device work is done."""
import ctypes

import numpy as np
import pytest

from ip_cu_lbm_amd import _abi, cases


def _tables(case, which, dtype):
    lib = _abi.load_library()
    cfg, ini, keep = case.to_c()
    n = ctypes.c_int64()
    assert lib.ipl_build_tables(ctypes.byref(cfg), which, None, ctypes.byref(n)) == 0
    out = np.empty(n.value, dtype=dtype)
    assert lib.ipl_build_tables(ctypes.byref(cfg), which, out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(n)) == 0
    return out


def _box3d(q, nx=6, ny=5, nz=4, periodic=(0, 0, 0), flux=_abi.FLUX_REFERENCE, ratio=1.1, solid=None):
    px, wx = cases.axis_from_widths(cases.stretched_widths(nx, ratio))
    py, wy = cases.axis_from_widths(cases.stretched_widths(ny, ratio))
    pz, wz = cases.axis_from_widths(np.ones(nz))
    return cases.Case(name="box", lattice=q, n=(nx, ny, nz), pos=[px, py, pz], width=[wx, wy, wz],
                      periodic=periodic, flux_axes=flux, solid=solid)


CASES = {
    "tg2d": lambda: cases.lattice_taylor_green(nx=16, nz=8),
    "tg2d_walls": lambda: cases.lattice_taylor_green(nx=16, nz=8).replace(periodic=(0, 0, 0)),
    "q19_box": lambda: _box3d(19),
    "q27_box": lambda: _box3d(27),
    "q27_periodic_y": lambda: _box3d(27, periodic=(0, 1, 0)),
    "q19_zflux": lambda: cases.lock_exchange(nx=8, ny=4, nz=8),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_builder_tables_match_oracle(oracle_lib, name):
    case = CASES[name]()
    o = oracle_lib.Oracle(case)
    q, N, G = case.lattice, o.n_calc, o.n_ghost
    nt = N + G
    neib_ref = o.table(0).reshape(nt, q).T  # oracle is [cell][k]
    neib = _tables(case, 0, np.int32).reshape(q, nt)
    assert np.array_equal(neib, neib_ref)
    nb_ref = o.table(1).reshape(G, q).T
    nk_ref = o.table(2).reshape(G, q).T
    nb = _tables(case, 1, np.int32).reshape(q, G)
    nk = _tables(case, 2, np.int32).reshape(q, G)
    # entries k >= 1 are the ones the slip kernel reads
    assert np.array_equal(nb[1:], nb_ref[1:])
    assert np.array_equal(nk[1:], nk_ref[1:])


def test_builder_rejects_narrow_cells():
    case = cases.lattice_taylor_green(nx=16, nz=8)
    w = case.width[0].copy()
    w[3] = 0.5
    case.width[0] = w
    lib = _abi.load_library()
    cfg, ini, keep = case.to_c()
    n = ctypes.c_int64()
    assert lib.ipl_build_tables(ctypes.byref(cfg), 0, None, ctypes.byref(n)) == _abi.IPL_ERR_INVALID_CONFIG


def test_corner_scheme_rejects_the_3_point_stencil(oracle_lib):
    """Flux along x and y in 3D takes the corner scheme, whose tensor stencil is 5 x 5: a
    3-point stencil is refused by the library and by the oracle alike."""
    case = _box3d(19, nx=8, ny=8, nz=6).replace(stencil=3)
    lib = _abi.load_library()
    cfg, ini, keep = case.to_c()
    n = ctypes.c_int64()
    assert lib.ipl_build_tables(ctypes.byref(cfg), 0, None, ctypes.byref(n)) == _abi.IPL_ERR_INVALID_CONFIG
    assert lib.ipl_build_tables(ctypes.byref(case.replace(stencil=5).to_c()[0]), 0, None, ctypes.byref(n)) == _abi.IPL_OK
    with pytest.raises(ValueError, match="5-point"):
        oracle_lib.Oracle(case)


def test_stencil_weights_partition_of_unity():
    case = cases.lattice_taylor_green(nx=32, nz=8)
    w = _tables(case, 3, np.float64).reshape(3, 2, 5, -1)
    s = w[0].sum(axis=1)
    assert np.abs(s - 1).max() < 1e-13
    assert np.abs(w[1]).max() == 0 and np.abs(w[2]).max() == 0


def test_library_exports_every_declared_symbol():
    import os
    import re
    hdr = open(os.path.join(os.path.dirname(__file__), "..", "include", "ipl.h")).read()
    names = set(re.findall(r"^\s*(?:int|const char\*)\s+(ipl_\w+)\s*\(", hdr, re.M))
    assert len(names) >= 12
    lib = _abi.load_library()
    for n in names:
        assert hasattr(lib, n), n


def test_subdivided_axis_matches_driver_formula():
    """subdivided_axis reproduces the drivers' refined columns (SphereFall.cu:80-128): centres at
    (i + (2d+1)/(2 divx[i])) dx ratiox, widths dx ratiox / divx[i], contiguous cells, total length kept."""
    coarse = 0.013  # column centres 0.0065, 0.0195, ..., 0.0975, 0.1105: none on a bound
    divs = cases.sphere_fall_divs(9, coarse)
    assert divs == [1, 1, 4, 4, 4, 4, 4, 2, 1]
    pos, wid = cases.subdivided_axis(divs, coarse)
    assert len(pos) == sum(divs)
    k = 0
    for i, d in enumerate(divs):
        for dv in range(d):
            assert abs(pos[k] - (i + (2.0 * dv + 1.0) / (2.0 * d)) * coarse) < 1e-15
            assert abs(wid[k] - coarse / d) < 1e-18
            k += 1
    edges = pos - wid / 2
    assert np.allclose(edges[1:], (pos + wid / 2)[:-1], atol=1e-15) and abs(wid.sum() - 9 * coarse) < 1e-14


def test_refined_axis_builds_tables():
    """A refined x axis (flux axis, dz = finest width) builds; the stencil weights of every cell are
    a partition of unity."""
    px, wx = cases.subdivided_axis([1, 2, 4, 4, 2, 1, 1, 1], 4.0)
    pz, wz = cases.axis_from_widths(np.ones(6))
    case = cases.Case(name="refined", lattice=9, n=(len(px), 1, 6), pos=[px, np.array([0.5]), pz],
                      width=[wx, np.array([1.0]), wz], periodic=(1, 0, 0), flux_axes=_abi.FLUX_X)
    w = _tables(case, 3, np.float64).reshape(3, 2, 5, -1)
    assert np.abs(w[0].sum(axis=1) - 1).max() < 1e-13


def test_read_initial_csv_reference_files():
    """The reference's input/initial.csv of three drivers (fixtures copied from LockExchange,
    Horn2001S and Boegman2005run18IB) through input_items + setc(9) (basic.cu:57-77): grid
    reduction by the ratios, the lattice choice, the step count and the relaxation times."""
    import os
    from ip_cu_lbm_amd import cases as C
    g = os.path.join(os.path.dirname(__file__), "golden")
    le = C.read_initial_csv(os.path.join(g, "initial_LockExchange.csv"))
    assert (le.nx, le.ny, le.nz, le.num_velocity) == (800, 1, 100, 9)
    assert le.total_count == 1 + 1000 * 30 and le.save_interval == 1000
    assert le.c == 1.0 and abs(le.nu - 1e-6) < 1e-21 and le.gz == -9.81
    assert abs(le.tau - (3e-6 / 0.001 + 0.5)) < 1e-15            # 0.503
    assert abs(le.taus - (1.4 * 3e-9 / 0.001 + 0.5)) < 1e-15
    horn = C.read_initial_csv(os.path.join(g, "initial_Horn2001S.csv"))
    assert (horn.nx, horn.ny, horn.nz, horn.num_velocity, horn.ratioy) == (200, 1, 58, 9, 1.0)  # 1200/6; 1/12 -> 1
    assert abs(horn.c - 2.0) < 1e-15 and abs(horn.tau - (3e-6 / 4.0 / 0.0025 + 0.5)) < 1e-15
    boeg = C.read_initial_csv(os.path.join(g, "initial_Boegman2005run18IB.csv"))
    assert (boeg.nx, boeg.nz) == (50, 145) and abs(boeg.c - 2.0) < 1e-15
    f = boeg.case_fields()
    assert f["lattice"] == 9 and f["n"] == (50, 1, 145) and f["dz"] == 0.002 and f["steps"] == 30001


def test_read_rates_csv_reference_files():
    """S_D3Q9.csv / S_D3Q27.csv of the reference (fixtures) as relaxation rates, stress rows zeroed
    as set_M does (basic.cu:184-185)."""
    import os
    import numpy as np
    from ip_cu_lbm_amd import cases as C
    g = os.path.join(os.path.dirname(__file__), "golden")
    raw9 = np.array([float(x) for x in open(os.path.join(g, "S_D3Q9.csv")).readline().split(",")[:9]])
    r9 = C.read_rates_csv(os.path.join(g, "S_D3Q9.csv"), 9)
    assert np.array_equal(r9[:7], raw9[:7]) and r9[7] == r9[8] == 0.0
    r27 = C.read_rates_csv(os.path.join(g, "S_D3Q27.csv"), 27)
    assert r27.shape == (27,) and np.all(r27[5:10] == 0.0)
