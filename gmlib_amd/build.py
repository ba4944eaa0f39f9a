"""Build recipe of libgml.so (HIP for gfx950) and the CPU oracle.

Usage: python -m gmlib_amd.build   (or __graft_entry__.build())
       python -m gmlib_amd.build --asan   (host-AddressSanitizer build in
       gmlib_amd/_build_asan/, for the CPU tests:
       GMLIB_AMD_LIB=gmlib_amd/_build_asan/libgml.so
       LD_PRELOAD=<clang>/lib/linux/libclang_rt.asan-x86_64.so
       ASAN_OPTIONS=detect_leaks=0 python -m pytest tests -m "not gpu")
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
    "gml_plan_gpu.hip",
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


# host code with AddressSanitizer (device code unchanged): a checking build
# for the CPU tests, run with GMLIB_AMD_LIB pointing at it and the
# sanitizer's runtime preloaded
ASAN = ["-g", "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer"]
ASAN_LIB = os.path.join(HERE, "_build_asan", "libgml.so")


def build_lib(verbose=False, asan=False):
    bdir = os.path.dirname(ASAN_LIB) if asan else BUILD
    lib = ASAN_LIB if asan else LIB
    os.makedirs(bdir, exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(bdir, src + ".o")
        cmd = [HIPCC] + COMMON + DEVICE + (ASAN if asan else []) + ["-x", "hip", "-c", path, "-o", obj]
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
    tmp = lib + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH] + (ASAN[1:3] if asan else []) + ["-o", tmp] + objs +
         ["-L/opt/rocm/lib", "-lhiprtc", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib", "-pthread"])
    os.replace(tmp, lib)
    return lib


def build_oracle():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def main():
    if "--asan" in sys.argv:
        print(build_lib(verbose=True, asan=True))
        return
    build_oracle()
    print(build_lib(verbose=True))


if __name__ == "__main__":
    main()
