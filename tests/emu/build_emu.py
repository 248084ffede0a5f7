"""Builds the CPU emulation of the HIP library (test infrastructure): every source of
ip_cu_lbm_amd/csrc compiled by g++ against tests/emu/hip_emu.hpp, optionally with
AddressSanitizer.  Output: tests/emu/_build/libipl_emu[_asan].so, same C ABI as libipl.so."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "ip_cu_lbm_amd", "csrc")
OUT = os.path.join(HERE, "_build")


def build(asan: bool = False) -> str:
    os.makedirs(OUT, exist_ok=True)
    lib = os.path.join(OUT, "libipl_emu_asan.so" if asan else "libipl_emu.so")
    srcs = [os.path.join(CSRC, f) for f in sorted(os.listdir(CSRC)) if f.endswith((".hip", ".cpp"))]
    srcs.append(os.path.join(HERE, "emu_runtime.cpp"))
    deps = srcs + [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hpp")] + [os.path.join(HERE, "hip_emu.hpp")]
    if os.path.exists(lib) and os.path.getmtime(lib) >= max(os.path.getmtime(d) for d in deps):
        return lib
    flags = ["-std=c++20", "-fPIC", "-shared", "-pthread", "-DIPL_CPU_EMULATION", "-include", os.path.join(HERE, "hip_emu.hpp"),
             "-I", os.path.join(ROOT, "include"), "-w"]
    flags += ["-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer"] if asan else ["-O2"]
    cmd = ["g++"] + flags + ["-x", "c++"] + srcs + ["-o", lib + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stderr[-8000:])
    os.replace(lib + ".tmp", lib)
    return lib


if __name__ == "__main__":
    print(build(asan="--asan" in sys.argv))
