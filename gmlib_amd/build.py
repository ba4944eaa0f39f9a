"""Build recipe of libgml.so (HIP for gfx950) and the CPU oracle.

Usage: python -m gmlib_amd.build   (or __graft_entry__.build())
Every HIP translation unit is compiled with hipcc for gfx950 only, with
packed-FP32 instructions disabled, then linked with hipRTC into one shared
library next to this file.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgml.so")

SOURCES = [
    "gml_runtime.cpp",
    "gml_codegen.cpp",
    "gml_extract.cpp",
    "gml_meshb.cpp",
    "gml_meshgen.cpp",
    "gml_topology.hip",
    "gml_reduce.hip",
    "gml_algebra.hip",
    "gml_mesh_kernels.hip",
    "gml_plan.cpp",
    "gml_batch.hip",
    "gml_laplacian.cpp",
    "gml_serve.cpp",
    "gml_renumber.hip",
    "gml_items.hip",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", "-Wno-unused-result",
          "-I" + CSRC, "-I" + os.path.join(ROOT, "include")]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = r.stdout
    # hipcc's host pass does not know the packed-fp32 feature: harmless
    out = "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l or "error" in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("command failed: " + " ".join(cmd))
    return out


def build_lib(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(BUILD, src + ".o")
        cmd = [HIPCC] + COMMON + DEVICE + ["-x", "hip", "-c", path, "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    failed = []
    for cmd, p in procs:
        out, _ = p.communicate()
        lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l]
        if p.returncode != 0:
            failed.append((cmd, "\n".join(lines)))
        elif verbose and lines:
            print("\n".join(lines))
    if failed:
        for cmd, out in failed:
            sys.stderr.write(" ".join(cmd) + "\n" + out + "\n")
        raise RuntimeError("HIP compilation failed")
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs +
         ["-L/opt/rocm/lib", "-lhiprtc", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib", "-pthread"])
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def main():
    build_oracle()
    print(build_lib(verbose=True))


if __name__ == "__main__":
    main()
