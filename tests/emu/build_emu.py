"""Builds the CPU emulation of the HIP library (test infrastructure): every source of
ip_cu_lbm_amd/csrc, plus the FP16 instantiations that only this build carries, compiled by g++
against tests/emu/hip_emu.hpp, optionally with AddressSanitizer.  Output:
tests/emu/_build/libipl_emu[_asan].so, the same C ABI as libipl.so.  Sources compile in parallel,
each to its own object, and only when it or a header changed."""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "ip_cu_lbm_amd", "csrc")
OUT = os.path.join(HERE, "_build")


def _sources():
    srcs = [os.path.join(CSRC, f) for f in sorted(os.listdir(CSRC)) if f.endswith((".hip", ".cpp"))]
    srcs += [os.path.join(HERE, "emu_runtime.cpp")]
    srcs += [os.path.join(HERE, f) for f in sorted(os.listdir(HERE)) if f.startswith("fused_f16_emu") and f.endswith(".cpp")]
    return srcs


def _headers():
    hs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hpp")]
    return hs + [os.path.join(HERE, "hip_emu.hpp"), os.path.join(ROOT, "include", "ipl.h")]


def build(asan: bool = False, jobs: int = 8) -> str:
    tag = "asan" if asan else "opt"
    obj_dir = os.path.join(OUT, "obj_" + tag)
    os.makedirs(obj_dir, exist_ok=True)
    lib = os.path.join(OUT, "libipl_emu_asan.so" if asan else "libipl_emu.so")
    flags = ["-std=c++20", "-fPIC", "-pthread", "-DIPL_CPU_EMULATION", "-include", os.path.join(HERE, "hip_emu.hpp"),
             "-I", os.path.join(ROOT, "include"), "-w"]
    flags += ["-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer"] if asan else ["-O2"]
    newest_header = max(os.path.getmtime(h) for h in _headers())

    def compile_one(src):
        obj = os.path.join(obj_dir, os.path.basename(src) + ".o")
        if os.path.exists(obj) and os.path.getmtime(obj) >= max(newest_header, os.path.getmtime(src)):
            return obj
        r = subprocess.run(["g++"] + flags + ["-x", "c++", "-c", src, "-o", obj + ".tmp"], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{os.path.basename(src)}:\n{r.stderr[-8000:]}")
        os.replace(obj + ".tmp", obj)
        return obj

    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, _sources()))
    if os.path.exists(lib) and os.path.getmtime(lib) >= max(os.path.getmtime(o) for o in objs):
        return lib
    link = ["g++", "-shared", "-pthread"] + (["-fsanitize=address"] if asan else []) + objs + ["-o", lib + ".tmp"]
    r = subprocess.run(link, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stderr[-8000:])
    os.replace(lib + ".tmp", lib)
    return lib


if __name__ == "__main__":
    print(build(asan="--asan" in sys.argv))
