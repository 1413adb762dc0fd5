"""Builds libbqcd_hip.so (the gfx950 BQCD hot-path library) in-tree.

    python bqcd-r399-quda_amd/build_lib.py          # compile + link
    python bqcd-r399-quda_amd/build_lib.py --check  # report register usage

Every HIP translation unit is compiled by hipcc for gfx950 with packed FP32
instructions disabled; objects go to _build/ next to this file and the
shared library to libbqcd_hip.so beside it.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libbqcd_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["hb_context.cpp", "hb_fields.hip", "hb_dslash.hip", "hb_blas.hip", "hb_cg.hip",
           "hb_gauge.hip", "hb_mixed.hip", "hb_multi.hip", "hb_noalloc.hip", "hb_plugin.hip",
           "hb_force.hip", "hb_hmc.hip", "hb_conf.cpp"]

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    # packed FP32 ops disabled in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _compile(src, extra):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + extra + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (p.stdout + p.stderr).splitlines() if "packed-fp32-ops" not in l)
    return src, obj, p.returncode, msgs


def build(check=False, jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    extra = ["-Rpass-analysis=kernel-resource-usage"] if check else []
    objs, failed = [], []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, obj, rc, msgs in ex.map(lambda s: _compile(s, extra), SOURCES):
            if rc != 0:
                failed.append((src, msgs))
            elif msgs.strip() and (check or "warning" in msgs):
                print("[%s]\n%s" % (src, msgs))
            objs.append(obj)
    if failed:
        for src, msgs in failed:
            print("[%s] FAILED\n%s" % (src, msgs), file=sys.stderr)
        raise RuntimeError("HIP compilation failed: " + ", ".join(s for s, _ in failed))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(check="--check" in sys.argv))
