"""CPU baseline of the generated-kernel path: measurement and test
infrastructure, never part of the product (the library never loads what is
built here, and nothing here runs on the GPU).

The reference runs its generated OpenCL kernels on whatever OpenCL device it
finds (sources/gmlib3.c:2502-2938 writes each kernel: the user's OpenCL-C body
between a generated prologue of index and memory reads and an epilogue of
writes).  This image has no CPU OpenCL runtime, so the reference-style CPU
path is rebuilt here from the engine's own generated kernels: the engine
writes a kernel for a host-only handle in the one-row launch form without LDS
staging (the OpenCL-C prelude, the GMlib prologue/epilogue and the
project-written OpenCL-C body, exactly the text hipRTC compiles for the GPU),
gml_host_shim.h stands in for the few HIP builtins that text uses, clang
compiles it for x86-64, and a generated loop runs every work-item of every
work-group over the host cores with OpenMP (work-groups split statically
over the threads, the work-items of one group in order on one thread).

The smoothing set below is the Laplacian smoothing sweep of the headline
workload (kernels/smooth_scatter.cl, kernels/smooth_gather.cl); its inputs
use the generated kernels' argument layouts: tet rows as int4, coordinates
and barycentres as float4, the vertex -> tet ball as 64-vertex slices of
column-major rows (sell64 below), the vertex references, the parameter block.
"""
import ctypes as C
import hashlib
import os
import re
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_host")
CLANG = os.environ.get("GML_HOST_CLANG", "/opt/rocm/llvm/bin/clang++")
OMP_LIB = "/opt/rocm/llvm/lib"
FLAGS = ["-std=c++17", "-O3", "-march=x86-64-v3", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
         "-fno-fast-math", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
         "-Wl,-rpath," + OMP_LIB]

_SIG = re.compile(r'extern "C" __global__ void __launch_bounds__\(\d+\) (\w+)\((.*?)\)\s*\{', re.S)


def kernel_signature(src):
    """Name and (type, name) arguments of a generated kernel's text."""
    m = _SIG.search(src)
    if not m:
        raise ValueError("no generated kernel header found")
    args = []
    for a in m.group(2).split(","):
        a = " ".join(a.split())
        n = re.search(r"(\w+)$", a).group(1)
        args.append((a[:len(a) - len(n)].strip(), n))
    return m.group(1), args


def wrapper(src):
    """The work-item loop of one kernel: extern "C" gml_host_<name>(pointers,
    n rows, first row, work-group size), OpenMP over the work-groups."""
    body = src.split("// KERNEL HEADER")[1]
    if "__shared__" in body or "__syncthreads" in body or "__local" in body or "barrier(" in body:
        raise ValueError("the host loop runs the one-row form without LDS staging only")
    name, args = kernel_signature(src)
    call, k = [], 0
    for t, n in args:
        if "*" in t:
            call.append("(%s)a[%d]" % (t, k))
            k += 1
        elif n == "count":
            call.append("(int2){n, first}")
        elif n == "gml_remap":
            call.append("0")
        else:
            raise ValueError("unexpected scalar argument %s %s" % (t, n))
    return name, k, """
#include <omp.h>
extern "C" int gml_host_%(name)s(void *const *a, int n, int first, int grp)
{
   const long nb = ((long)n + grp - 1) / grp;
#pragma omp parallel for schedule(static)
   for(long b = 0; b < nb; b++)
   {
      blockIdx = {(unsigned int)b, 0, 0};
      blockDim = {(unsigned int)grp, 1, 1};
      gridDim = {(unsigned int)nb, 1, 1};
      for(int t = 0; t < grp; t++)
      {
         threadIdx = {(unsigned int)t, 0, 0};
         %(name)s(%(call)s);
      }
   }
   return %(nptr)d;
}
""" % {"name": name, "call": ", ".join(call), "nptr": k}


def digest(sources):
    """Identifies a build: the kernel texts, the shim and the compiler line."""
    parts = [sources[k] for k in sorted(sources)] + [open(os.path.join(HERE, "gml_host_shim.h")).read(),
                                                     CLANG] + FLAGS
    return hashlib.sha256("\0".join(parts).encode()).hexdigest()[:16]


def build(sources, out=OUT, tag="smoothing"):
    """Compile {kernel name: generated text} for the host into
    <out>/libgml_host_<tag>.so (one translation unit per kernel).  Rebuilds
    only when the texts changed.  Returns the library path."""
    os.makedirs(out, exist_ok=True)
    so = os.path.join(out, "libgml_host_%s.so" % tag)
    stamp = os.path.join(out, "libgml_host_%s.sha" % tag)
    h = digest(sources)
    if os.path.exists(so) and os.path.exists(stamp) and open(stamp).read().strip() == h:
        return so
    units = []
    for i, k in enumerate(sorted(sources)):
        name, _, wrap = wrapper(sources[k])
        if name != k:
            raise ValueError("kernel %s is named %s in its text" % (k, name))
        cpp = os.path.join(out, "%s_%s.cpp" % (tag, k))
        with open(cpp, "w") as f:
            f.write('#include "%s"\n' % os.path.join(HERE, "gml_host_shim.h"))
            f.write(sources[k])
            f.write(wrap)
            if i == 0:
                f.write('extern "C" int gml_host_threads() { return omp_get_max_threads(); }\n')
        units.append(cpp)
    tmp = so + ".tmp"
    r = subprocess.run([CLANG] + FLAGS + ["-o", tmp] + units, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("host build of the generated kernels failed:\n" + r.stdout[-4000:])
    os.replace(tmp, so)
    with open(stamp, "w") as f:
        f.write(h + "\n")
    return so


class HostKernels:
    """The built library of the texts `sources`: run(name, arrays, n) runs
    kernel `name` over rows first .. first+n-1 on every host thread, its
    pointer arguments taken by name from `arrays` (numpy arrays)."""

    def __init__(self, so, sources):
        self.lib = C.CDLL(so)
        self.lib.gml_host_threads.restype = C.c_int
        self.threads = self.lib.gml_host_threads()
        self.args = {k: [n for t, n in kernel_signature(s)[1] if "*" in t] for k, s in sources.items()}

    def run(self, name, arrays, n, first=0, grp=256):
        f = getattr(self.lib, "gml_host_" + name)
        f.restype = C.c_int
        f.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_int]
        ptrs = [arrays[a].ctypes.data for a in self.args[name]]
        if f((C.c_void_p * len(ptrs))(*ptrs), n, first, grp) != len(ptrs):
            raise RuntimeError("host kernel %s: argument count mismatch" % name)


def sell64(deg, ball):
    """A dense vertex -> tet ball (V x w packed entries, deg per vertex) in the
    generated gather's layout: vertex v's entry k at
    off[v >> 6] * 64 + k * 64 + (v & 63), one 64-vertex slice per offset, each
    slice as wide as its largest degree.  Returns (table, offsets)."""
    V = len(deg)
    nc = (V + 63) // 64
    d = np.zeros(nc * 64, np.int64)
    d[:V] = deg
    w = d.reshape(nc, 64).max(1)
    off = np.zeros(nc, np.int64)
    off[1:] = np.cumsum(w)[:-1]
    tab = np.zeros(max(int(w.sum()), 1) * 64, np.int32)
    v = np.arange(V, dtype=np.int64)
    base = off[v >> 6] * 64 + (v & 63)
    for k in range(ball.shape[1]):
        m = k < deg
        if not m.any():
            break
        tab[base[m] + k * 64] = ball[m, k]
    return tab, off.astype(np.uint32)


SMOOTH_KERNELS = ("smooth_scatter", "smooth_gather")


def smoothing_sources(G, kernel_dir, omega_block=16, n=4):
    """The generated texts of the smoothing sweep, written by a host-only
    handle of the engine (module G = gmlib_amd) for an n^3 Kuhn cube, one-row
    form, LDS staging off.  The ball width (the gather's BaryDegMax) follows
    the mesh's largest vertex degree, which is the same for every Kuhn cube
    with n >= 2."""
    crd, ref, tet = G.cube_mesh(n, 0.2, 3, True)
    body = lambda k: open(os.path.join(kernel_dir, k + ".cl")).read()
    with G.Gml(host_only=True) as g:
        G.lib().GmlxSetStaging(g.h, 0)
        vi, ti = g.set_mesh(crd, ref, tet)
        g.new_parameters(omega_block, body("smooth_params"))
        b = g.solution(G.GmlTetrahedra, 1, G.GmlFlt4, "Bary")
        ks = g.compile(body("smooth_scatter"), "smooth_scatter", G.GmlTetrahedra,
                       [(vi, G.GmlReadMode, 0), (b, G.GmlWriteMode, 0)])
        kg = g.compile(body("smooth_gather"), "smooth_gather", G.GmlVertices,
                       [(b, G.GmlReadMode, 0), (vi, G.GmlReadMode | G.GmlWriteMode | G.GmlRefFlag, 0)])
        return {"smooth_scatter": g.L.GmlxGetKernelSource(g.h, ks).decode(),
                "smooth_gather": g.L.GmlxGetKernelSource(g.h, kg).decode()}


def quality_sources(G, kernel_dir, n=4):
    """The generated texts of the 20^3 quality step's kernels (the tet
    mean-ratio body over the tet -> vertex downlink, the ball minimum body
    over the vertex -> tet ball), written as smoothing_sources does."""
    crd, ref, tet = G.cube_mesh(n, 0.2, 3, True)
    body = lambda k: open(os.path.join(kernel_dir, k + ".cl")).read()
    with G.Gml(host_only=True) as g:
        G.lib().GmlxSetStaging(g.h, 0)
        vi, ti = g.set_mesh(crd, ref, tet)
        q = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "qal")
        vq = g.solution(G.GmlVertices, 1, G.GmlFlt, "vq")
        kq = g.compile(body("tet_meanratio"), "tet_meanratio", G.GmlTetrahedra,
                       [(vi, G.GmlReadMode, 0), (q, G.GmlWriteMode, 0)])
        kb = g.compile(body("ball_min"), "ball_min", G.GmlVertices, [(q, G.GmlReadMode, 0), (vq, G.GmlWriteMode, 0)])
        return {"tet_meanratio": g.L.GmlxGetKernelSource(g.h, kq).decode(),
                "ball_min": g.L.GmlxGetKernelSource(g.h, kb).decode()}


def ball_width(src, name="Bary"):
    m = re.search(r"#define\s+%sDegMax\s+(\d+)" % name, src)
    return int(m.group(1)) if m else None


class Smoothing:
    """Jacobi sweeps of the generated smoothing kernels on the host over one
    mesh: crd (V x 4 float32, updated in place), ref, tet, the dense ball
    (deg, V x w packed entries) and omega."""

    def __init__(self, hk, crd, ref, tet, deg, ball, omega):
        self.hk = hk
        self.V, self.T = len(crd), len(tet)
        self.crd = crd
        self.tet = np.ascontiguousarray(tet, np.int32)
        self.ref = np.ascontiguousarray(ref, np.int32)
        self.deg = np.ascontiguousarray(deg, np.int32)
        self.tab, self.off = sell64(deg, ball)
        self.bary = np.zeros((self.T, 4), np.float32)
        self.par = np.zeros(4, np.float32)
        self.par[0] = omega
        self.arrays = {"TetVerTab": self.tet, "VerCrdTab": self.crd, "BaryTab": self.bary, "GmlPar": self.par,
                       "VerTetBalTab": self.tab, "VerTetBalOff": self.off, "VerTetDegTab": self.deg,
                       "VerRefTab": self.ref}

    def sweep(self):
        self.hk.run("smooth_scatter", self.arrays, self.T)
        self.hk.run("smooth_gather", self.arrays, self.V)


class Quality:
    """The generated quality and ball-minimum kernels on the host over one
    mesh: crd (V x 4 float32), tet, the dense ball (deg, V x w packed
    entries).  step() fills qal (per tet) and vq (per vertex)."""

    def __init__(self, hk, crd, tet, deg, ball):
        self.hk = hk
        self.V, self.T = len(crd), len(tet)
        self.crd = np.ascontiguousarray(crd, np.float32)
        self.tet = np.ascontiguousarray(tet, np.int32)
        self.deg = np.ascontiguousarray(deg, np.int32)
        self.tab, self.off = sell64(deg, ball)
        self.qal = np.zeros(self.T, np.float32)
        self.vq = np.zeros(self.V, np.float32)
        self.par = np.zeros(4, np.float32)
        self.arrays = {"TetVerTab": self.tet, "VerCrdTab": self.crd, "qalTab": self.qal, "vqTab": self.vq,
                       "GmlPar": self.par, "VerTetBalTab": self.tab, "VerTetBalOff": self.off,
                       "VerTetDegTab": self.deg}

    def step(self):
        self.hk.run("tet_meanratio", self.arrays, self.T)
        self.hk.run("ball_min", self.arrays, self.V)
