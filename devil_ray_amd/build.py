"""Builds libdrx.so (gfx950) in-tree: devil_ray_amd/libdrx.so.

Every .hip source is compiled by hipcc for gfx950 only, with packed-FP32 instruction selection
disabled; drx_lbvh.hip additionally with -ffp-contract=off so the BVH build rounds exactly like the
float32 CPU path it mirrors.  Objects go to devil_ray_amd/_obj/.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libdrx.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = {
    "drx_runtime.hip": [],
    "drx_lbvh.hip": ["-ffp-contract=off"],
    "drx_query.hip": [],
    "drx_integrate.hip": [],
    "drx_wide.hip": [],
    "drx_mlp.hip": [],
    "drx_image.hip": [],
    "dray_host.hip": [],
    "dray_blueprint.hip": [],
}
ROOT = os.path.dirname(HERE)
CPP_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "dray_api_main.cpp")
CPP_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "dray_api_test")
CPP_BP_SRC = os.path.join(ROOT, "tests", "cpp", "dray_blueprint_main.cpp")
CPP_BP_BIN = os.path.join(ROOT, "tests", "cpp", "dray_blueprint_test")

COMMON = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose: bool = False, defines=(), out: str = LIB, obj_dir: str = OBJ) -> str:
    """Compile and link; `defines` are extra -D flags (compile-time tuning knobs), `out` the .so path."""
    os.makedirs(obj_dir, exist_ok=True)
    procs = []
    objs = []
    for src, extra in SOURCES.items():
        obj = os.path.join(obj_dir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC] + COMMON + extra + [f"-D{d}" for d in defines] + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for src, cmd, p in procs:
        outb, _ = p.communicate()
        text = outb.decode(errors="replace")
        if p.returncode != 0:
            failed.append((src, text))
        elif verbose and text.strip():
            print(text)
    if failed:
        msg = "\n".join(f"--- {s} ---\n{t}" for s, t in failed)
        raise RuntimeError(f"hipcc failed:\n{msg}")
    tmp = out + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.check_call(cmd)
    shutil.move(tmp, out)
    return out


def build_cpp_test(lib: str = LIB, out: str = CPP_TEST_BIN) -> str:
    """The C++ programs that drive the path through the dray-named classes (linked against libdrx.so):
    tests/cpp/dray_api_test and tests/cpp/dray_blueprint_test (beside `out`)."""
    rpath = os.path.relpath(os.path.dirname(lib), os.path.dirname(out))
    for src, dst in ((CPP_TEST_SRC, out), (CPP_BP_SRC, os.path.join(os.path.dirname(out), os.path.basename(CPP_BP_BIN)))):
        tmp = dst + ".tmp"
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), src, "-o", tmp,
               "-L", os.path.dirname(lib), "-l:" + os.path.basename(lib), f"-Wl,-rpath,$ORIGIN/{rpath}",
               "-Wl,-rpath-link,/opt/rocm/lib"]
        subprocess.check_call(cmd)
        shutil.move(tmp, dst)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_cpp_test())
