"""File writers on the GPU path: the reference's output() CSV (basic.cu:3-30), IB_csv
(ibm.cu:3-14) and the VTU of make_VTK.py, checked against the fields read back."""
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from ip_cu_lbm_amd import F64, PATH_FUSED, Simulation, _abi, cases

pytestmark = pytest.mark.gpu

FIELDS = ("p", "S", "phi", "rho", "u", "v", "w", "Fx", "Fy", "Fz")


def test_fields_csv_format(tmp_path):
    case = cases.lock_exchange(nx=12, ny=4, nz=8, path=PATH_FUSED)
    with Simulation(case) as sim:
        sim.step(5)
        vals = {f: sim.read(f) for f in FIELDS}
        path = tmp_path / "Fakhari_LBM0001.csv"
        sim.write_fields_csv(str(path))
    lines = path.read_text().splitlines()
    assert lines[0] == f"num_calc,{case.n_cells}"
    assert lines[1] == ("x       ,y       ,z       ,delx    ,dely    ,pressure,salinity ,phi     ,rho     ,"
                        "u       ,v       ,w       ,Fx      ,Fy      ,Fz      ")
    rows = np.array([[float(x) for x in l.split(",")] for l in lines[2:]])
    assert rows.shape == (case.n_cells, 15)
    X, Y, Z = case.coords()
    for col, want in ((0, X.ravel()), (1, Y.ravel()), (2, Z.ravel())):
        assert np.abs(rows[:, col] - want).max() <= 5e-7
    for col, f in enumerate(FIELDS, start=5):
        assert np.abs(rows[:, col] - vals[f]).max() <= 5e-7, f  # "%f": six decimals
    assert lines[2].count(",") == 14 and "." in lines[2].split(",")[5]


def test_vtu_cells_and_data(tmp_path):
    case = cases.lock_exchange(nx=10, ny=3, nz=7, path=PATH_FUSED)
    with Simulation(case) as sim:
        sim.step(3)
        vals = {f: sim.read(f) for f in FIELDS}
        path = tmp_path / "LBM0000.vtu"
        sim.write_vtu(str(path))
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    n = case.n_cells
    assert int(piece.get("NumberOfCells")) == n
    assert int(piece.get("NumberOfPoints")) == 11 * 4 * 8  # shared corners of a full box
    pts = np.array(piece.find("Points/DataArray").text.split(), dtype=float).reshape(-1, 3)
    arrays = {a.get("Name"): a for a in piece.find("Cells")}
    conn = np.array(arrays["connectivity"].text.split(), dtype=np.int64).reshape(n, 8)
    assert set(np.array(arrays["types"].text.split(), dtype=int)) == {12}
    # every hexahedron spans its cell: centre and widths from the corners
    X, Y, Z = case.coords()
    lo, hi = pts[conn[:, 0]], pts[conn[:, 6]]
    assert np.abs((lo + hi) / 2 - np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)).max() < 1e-12
    widths = np.stack(np.meshgrid(case.width[0], case.width[1], case.width[2], indexing="ij"), -1).reshape(-1, 3)
    assert np.abs((hi - lo) - widths).max() < 1e-12
    cd = {a.get("Name"): np.array(a.text.split(), dtype=float) for a in piece.find("CellData")}
    for name, f in (("salinity", "S"), ("phi", "phi"), ("rho", "rho"), ("pressure", "p")):
        assert np.array_equal(cd[name], vals[f]), name
    assert np.array_equal(cd["velocity"].reshape(n, 3), np.stack([vals["u"], vals["v"], vals["w"]], 1))
    assert np.array_equal(cd["force"].reshape(n, 3), np.stack([vals["Fx"], vals["Fy"], vals["Fz"]], 1))


def test_markers_csv(tmp_path):
    c = cases.falling_sphere(n=(20, 18, 24), diameter=8.0, marker_set="3", storage=F64, path=PATH_FUSED, tau=0.7,
                             gravity=2e-4, method=_abi.IB_STRESS, kernel=_abi.IB_KERNEL_REFERENCE)
    c.body.pos = (10.3, 8.8, 12.4)
    with Simulation(c) as sim:
        sim.step(4)
        pos, vel, gw = sim.markers("pos"), sim.markers("vel"), sim.markers("force")
        path = tmp_path / "ibm_0001.csv"
        sim.write_markers_csv(str(path))
    lines = path.read_text().splitlines()
    assert lines[0] == "x,y,z,velx,vely,velz,ax,ay,az"
    rows = np.array([[float(x) for x in l.split(",")] for l in lines[1:]])
    assert rows.shape == (len(c.body.markers), 9)
    want = np.concatenate([pos, vel, gw], axis=1)
    # std::ostream's default six significant digits
    assert np.all(np.abs(rows - want) <= 1e-5 * np.abs(want) + 1e-300)


def test_settling_history_cd_csv(tmp_path):
    """The falling-sphere driver's C_D.csv: the body's z velocity at every save interval
    (SphereFall.cu:424, 437), recorded from the library's FP64 body state and read back."""
    from ip_cu_lbm_amd import settling_history, write_cd_csv
    c = cases.falling_sphere(n=(20, 18, 30), diameter=8.0, marker_set="3", storage=F64, path=PATH_FUSED, tau=0.7,
                             gravity=2e-4, method=_abi.IB_STRESS, kernel=_abi.IB_KERNEL_REFERENCE, density_ratio=4.0)
    with Simulation(c) as sim:
        vel, t = settling_history(sim, 60, 20, c.dt)
        final = sim.body_state()["vel"][2]
    assert len(vel) == 3 and vel[-1] == final and vel[-1] < vel[0] < 0  # falling, accelerating
    assert np.allclose(t, [20 * c.dt, 40 * c.dt, 60 * c.dt])
    path = tmp_path / "C_D.csv"
    write_cd_csv(str(path), vel, t)
    rows = np.array([[float(x) for x in l.split(",")] for l in path.read_text().splitlines()[1:]])
    assert np.allclose(rows[:, 0], vel, rtol=1e-5) and np.allclose(rows[:, 1], t, rtol=1e-5)
