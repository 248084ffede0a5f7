"""The drivers' C_D.csv writer (out_C_D, SphereFall3D/SphereFall.cu:32-42; host only, no GPU):
header "C_D,C_time", one row per record, each value in the C++ stream's default format (%g, six
significant digits) as the reference writes it, e.g. its Karman_vortex_2D/C_D.csv."""
import os

import numpy as np

from ip_cu_lbm_amd import write_cd_csv

GOLDEN_CD = os.path.join(os.path.dirname(__file__), "golden", "C_D_Karman_vortex_2D_head.csv")


def test_cd_csv_format(tmp_path):
    vals = np.array([2.75106, -0.0175131234, 1e-7, 123456789.0, 0.0])
    times = np.array([9.9996, 19.9992, 29.9988, 39.9984, 0.5])
    path = tmp_path / "C_D.csv"
    write_cd_csv(str(path), vals, times)
    lines = path.read_text().splitlines()
    assert lines[0] == "C_D,C_time"
    assert lines[1:] == [f"{a:g},{t:g}" for a, t in zip(vals, times)]


def test_cd_csv_reproduces_the_reference_file_rows(tmp_path):
    """The first rows of the reference's Karman_vortex_2D/C_D.csv, parsed and written again,
    come out byte for byte."""
    ref = open(GOLDEN_CD).read().splitlines()
    rows = np.array([[float(x) for x in l.split(",")] for l in ref[1:]])
    path = tmp_path / "C_D.csv"
    write_cd_csv(str(path), rows[:, 0], rows[:, 1])
    assert path.read_text().splitlines() == ref


def test_cd_csv_empty(tmp_path):
    path = tmp_path / "C_D.csv"
    write_cd_csv(str(path), [], [])
    assert path.read_text() == "C_D,C_time\n"
