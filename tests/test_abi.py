"""The ctypes mirrors in ip_cu_lbm_amd/_abi.py match the C structs of include/ipl.h: size and the
offset of every field, as the host C compiler lays them out."""
import ctypes
import os
import subprocess

import pytest

from ip_cu_lbm_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _c_layout(tmp_path, struct, fields):
    src = tmp_path / "layout.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "ipl.h"', "int main(void) {",
             f'printf("size %zu\\n", sizeof({struct}));']
    for f in fields:
        lines.append(f'printf("{f} %zu\\n", offsetof({struct}, {f}));')
    lines += ["return 0;", "}"]
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    return {k: int(v) for k, v in (l.split() for l in out if l)}


@pytest.mark.parametrize("struct,cls", [("ipl_config", _abi.IplConfig), ("ipl_init", _abi.IplInit),
                                        ("ipl_ib_config", _abi.IplIbConfig)])
def test_struct_layout_matches_c(tmp_path, struct, cls):
    names = [f[0] for f in cls._fields_]
    c = _c_layout(tmp_path, struct, names)
    assert c["size"] == ctypes.sizeof(cls)
    for n in names:
        assert c[n] == getattr(cls, n).offset, n


def test_library_exports_every_declared_entry_point():
    """Every function include/ipl.h declares is exported by the built libipl.so (the library is
    loaded, not called: no device work)."""
    import re
    header = open(os.path.join(ROOT, "include", "ipl.h")).read()
    header = re.sub(r"/\*.*?\*/", "", header, flags=re.S)
    declared = set(re.findall(r"\b(ipl_[a-z0-9_]+)\s*\(", header))
    assert len(declared) > 30
    lib = _abi.load_library()
    missing = [name for name in sorted(declared) if not hasattr(lib, name)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", lib._name], check=True, capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.strip()}
    assert declared <= exported, sorted(declared - exported)
