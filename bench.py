"""Benchmark of the GMlib kernel engine on one MI355X.

Headline workload (one "step"): Laplacian node smoothing at scale -- a
256^3-cell Kuhn cube mesh (16,974,593 vertices, 100,663,296 tets, FP32
coordinates jittered by +-0.2h, seeded, boundary vertices fixed), 10 Jacobi
sweeps (tet barycentre scatter + vertex-ball gather), then a tet-quality
recompute and a GmlReduceVector min.  The reported path is the drop-in C API
(GmlCompileKernel / GmlLaunchKernel / GmlReduceVector over the generated
kernels of gmlib_amd/kernels/*.cl) with inputs resident in HBM.

Metric: `value` is the drop-in step's algorithmic HBM bytes (each datum
counted once: scatter 32T + 16V, gather 32T + 40V per sweep, quality
20T + 16V, min 4T -- the bytes this path moves at least) per second.  The
hand-written path (GmlxSmoothSweeps + GmlxTetQuality) is timed beside it
from the same initial coordinates, with its own bytes per second (its
operator-form sweep moves far fewer bytes) and, under model_equivalent_GBs,
the drop-in byte model divided by its step time -- a work rate, not an HBM
rate.  roofline: the dominant kernel's algorithmic bytes over its average
duration (HIP events on the library's stream) against the 8 TB/s peak.
parity: after the timed steps each path's coordinates against the oracle's
FP32 sweeps run from the same initial coordinates for as many sweeps, the
min quality bit-exact against the GPU's own quality array, the quality
against FP64 on the GPU's coordinates; that oracle run, timed on the host
cores, is the CPU baseline.

Second workload (--workload cg): block-sparse SpMV and block-Jacobi CG on the
3x3-block vector Laplacian of the edge graph of a 128^3-cell Kuhn cube
(2,146,689 block rows, 31.8M blocks), FP32 values from a seeded SPD edge
weighting.  A step is one GmlMultMatVec (y = A x), counted as 40 B per block
(36 B values + 4 B column) + 28 B per row (x, y, row length); the kernel's
own bytes (own_bytes_per_launch) count 2 B for a column it reads as a 16-bit
offset from its row.  Then a plain FP32 CG solve and an FP32 CG with FP64
residual refinement to 1e-10, each checked against the oracle's FP64 product
b - A x on the host.

Third workload (--workload batch): 512 Kuhn cube meshes with cells per side
drawn log-uniform in [8, 96] (seeded; jitter +-0.3h), kept on the device as
one batch.  A step is one GmlxBatchQuality: per-tet mean-ratio quality,
per-vertex ball minimum and per-mesh min / mean of every mesh, one
persistent kernel replayed from a HIP graph, launch shape picked by the
autotuner (GmlxBatchTune).  Counted as 36 B per tet (connectivity, 4 ball
slots, quality out) + 20 B per vertex (coordinates, ball minimum out).  The
plain per-mesh path (one handle per mesh; GmlxTetQuality, GmlxBallMin and two
GmlReduceVector per mesh) is timed beside it on the same meshes, and the
per-mesh results of both are compared.

Fourth workload (--workload serve): a served stream of mixed mesh jobs.  A
pool of seeded cube meshes with tet counts log-uniform in [10k, 20M]; open-
loop Poisson arrivals at a fixed rate, 50% quality evaluations, 30% 20-sweep
smoothing jobs, 20% block-CG solves to 1e-8 (FP64-refined), each on a mesh
drawn from the pool.  The server (GmlxServeStart) keeps meshes resident
under an HBM budget and runs up to --slots jobs at once; jobs/s and p50/p99
latency (submission to completion) are reported at that rate, beside the
saturated throughput (all jobs submitted at once) and the plain per-job path
(every job loads its mesh into a fresh handle, one at a time), whose results
must equal the served ones.

Fifth workload (--workload extract): GmlExtractEdges + GmlExtractFaces on the
256^3-cell Kuhn cube (100,663,296 tets) with the tables resident on the
device: every distinct edge and triangle once, in the reference's order
(first appearance), as new mesh tables.  Counted as the element table read
once per call (16 B per tet, twice) plus the tables written (edges 8 B nodes
+ 4 B ref, triangles 16 B + 4 B).  The CPU baseline is the oracle's
restatement of the reference's hash (one thread, as the reference) on a
smaller cube, in the same bytes per second.

Sixth workload (--workload heat): the reference's HeatCellCentered
application (gmlib_amd/heat.py) on a 128^3-cell Kuhn cube (12.6M tets,
25.3M faces): a step is one explicit iteration -- face values, cell
gradients, face gradients, flux balance, update (five generated kernels over
triangle -> tet shells and tet -> triangle downlinks) and the L2 residual,
counted as 120 B per tet + 72 B per face + 32 B per vertex.  The CPU
baseline is the numpy FP64 restatement of the same iteration on a smaller
cube.

Seventh workload (--workload bandwidth): the reference's GpuBandwidth
example -- four one-line bodies through the generated kernels: a vertex copy
(32 B per vertex), the tet indices as floats (32 B per tet), the indirect
tet -> vertex gather (16 B per vertex + 32 B per tet) and the ragged vertex
-> tet ball gather (32 B per tet + 20 B per vertex) -- on the 256^3 cube
numbered at random, then after GmlxHilbertRenumber; the example's own byte
figures are printed beside them.  The CPU baseline is the oracle's OpenMP
loops of the same four patterns on the renumbered mesh.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256]
                       [--workload all|smoothing|small|cg|batch|serve|extract|heat|bandwidth]
The default (all) measures the smoothing workload as the headline and the
20^3 quality, CG, batch and serve workloads as sub-records of the same line
("configs").  Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def body(name):
    return open(os.path.join(ROOT, "gmlib_amd", "kernels", name + ".cl")).read()


def new_handle(args):
    """A device handle on GPU 0 with the --staging choice applied."""
    import gmlib_amd as G
    g = G.Gml(0)
    if args.staging >= 0:
        g.L.GmlxSetStaging(g.h, args.staging)
    return g


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="all",
                    choices=["all", "smoothing", "small", "cg", "batch", "serve", "extract", "heat", "bandwidth"])
    ap.add_argument("--configs", default="1,3,3b,4,5",
                    help="all workload: the other configurations measured after the headline one")
    ap.add_argument("--qps", type=float, default=40.0, help="serve workload: arrival rate (jobs/s)")
    ap.add_argument("--jobs", type=int, default=400, help="serve workload: jobs in the timed stream")
    ap.add_argument("--pool", type=int, default=24, help="serve workload: meshes in the pool")
    ap.add_argument("--slots", type=int, default=8, help="serve workload: jobs running at once")
    ap.add_argument("--budget-gb", type=float, default=64.0, help="serve workload: HBM budget of resident meshes")
    ap.add_argument("--nmesh", type=int, default=512, help="batch workload: number of meshes")
    ap.add_argument("--plain", type=int, default=1, help="batch workload: also time the plain per-mesh path (0: skip)")
    ap.add_argument("--shape", default="", help="batch workload: TV/Pad to use instead of the autotuner (e.g. 256/4)")
    ap.add_argument("--staging", type=int, default=-1,
                    help="LDS staging of uplink gathers in generated kernels: 1 on, 0 off, -1 the library default")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256, help="cells per cube side")
    ap.add_argument("--cg-n", type=int, default=128, help="cg workload: cells per side of the cube whose edge graph "
                    "gives the matrix")
    ap.add_argument("--cg-solve", type=int, default=1, help="cg workload: also run the two CG solves (0: SpMV only)")
    ap.add_argument("--sweeps", type=int, default=10)
    ap.add_argument("--omega", type=float, default=0.2)
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="CPU baseline: run oracle sweeps for about this long (0: skip)")
    ap.add_argument("--sweep-tile", type=int, default=256, help="vertices per tile of the fused sweep")
    ap.add_argument("--quality-shape", default="", help="quality kernel launch: group/tets per thread (e.g. 256/2)")
    ap.add_argument("--paths", default="generated,builtin")
    ap.add_argument("--copy-calib", type=int, default=0,
                    help="launch a plain copy-like kernel of known size this many times first (smoothing: "
                         "float4 vertex copy, 16V read + 16V write; cg: GmlScaleVec, 12 B per row read and "
                         "written), to calibrate HBM byte counters")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: the path is single-device (multi-GPU sharding is a design "
              "note only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    runs = {"smoothing": run_smoothing, "small": run_small, "cg": run_cg, "batch": run_batch, "serve": run_serve,
            "extract": run_extract, "heat": run_heat, "bandwidth": run_bandwidth}
    if args.workload != "all":
        print(json.dumps(runs[args.workload](args)), flush=True)
        return
    # the headline configuration (256^3 smoothing), then each other
    # configuration as a sub-record of the same line, with a bounded CPU
    # baseline; a configuration that fails is recorded as such
    out = run_smoothing(args)
    sub = {}
    names = {"1": ("config1_quality_20cube", run_small), "3": ("config3_spmv_cg_128cube", run_cg),
             "3b": ("spmv_256cube", run_cg),
             "4": ("config4_batch_512meshes", run_batch), "5": ("config5_served_stream", run_serve)}
    for c in [x for x in args.configs.split(",") if x]:
        name, fn = names[c]
        a = argparse.Namespace(**vars(args))
        a.cpu_seconds = min(args.cpu_seconds, 3.0)
        if c == "3b":                   # the SpMV of the 256^3 cube's edge graph (100.7M tets), no solves
            a.cg_n, a.cg_solve = 256, 0
        t0 = time.time()
        log("---- %s ----" % name)
        try:
            rec = fn(a)
        except Exception as e:          # recorded, so the headline line still prints
            log("%s failed: %r" % (name, e))
            rec = {"error": repr(e)}
        for k in ("metric", "n_gpus", "higher_is_better", "scaling", "vs_baseline"):
            rec.pop(k, None)
        rec["wall_s"] = round(time.time() - t0, 1)
        sub[name] = rec
    out["configs"] = sub
    print(json.dumps(out), flush=True)


FORMS = {1: "one row", 2: "paired", 3: "element tiles"}


def tuned(L, h, k):
    """whether a generated kernel's launch-shape tuner has settled"""
    a, b = C.c_int(), C.c_int()
    L.GmlxGetKernelShape(h, k, C.byref(a), C.byref(b))
    return a.value != 0


def opencl_devices():
    """Device types the OpenCL platforms of this machine report (clinfo, in a
    child process): a CPU device would run the reference-style kernels
    directly; None if clinfo is absent."""
    import re
    import subprocess
    try:
        r = subprocess.run(["clinfo"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True, timeout=60)
    except (OSError, subprocess.TimeoutExpired):
        return None
    types = re.findall(r"Device Type:\s*(\S+)", r.stdout)
    return {"cpu": sum("CPU" in t for t in types), "gpu": sum("GPU" in t for t in types)}


def host_baseline(crd0, ref, tet, deg, ball, gpu_gather_src, omega, sweep_bytes, n):
    """CPU baseline of the smoothing sweep: the engine's own generated
    scatter and gather kernels (OpenCL-C bodies, prelude and GMlib
    prologue/epilogue, one-row form) compiled for x86-64 by clang and run over
    the host cores with OpenMP -- this image has no CPU OpenCL runtime to run
    them as the reference would.  One warm-up sweep, then timed sweeps; the
    value is the algorithmic sweep bytes over the median sweep time.
    Returns (record, coordinates after the sweeps, number of sweeps run)."""
    import numpy as np
    import gmlib_amd as G
    from oracle import host_kernels as HK

    t0 = time.time()
    src = HK.smoothing_sources(G, os.path.join(ROOT, "gmlib_amd", "kernels"))
    if HK.ball_width(src["smooth_gather"]) != HK.ball_width(gpu_gather_src):
        raise RuntimeError("host and device gathers differ in ball width")
    so = HK.build(src)
    hk = HK.HostKernels(so, src)
    c = np.array(crd0, copy=True)
    sm = HK.Smoothing(hk, c, ref, tet, deg, ball, omega)
    log("host build of the generated kernels + ball slices: %.1fs" % (time.time() - t0))
    sm.sweep()
    times = []
    while len(times) < 5 or (sum(times) < 8.0 and len(times) < 40):
        t1 = time.perf_counter()
        sm.sweep()
        times.append(time.perf_counter() - t1)
    med = float(np.median(times))
    rec = {"value": round(sweep_bytes / med / 1e9, 2), "unit": "GB/s", "cores": hk.threads, "kind": "port",
           "sweep_ms_median": round(med * 1e3, 2), "opencl_devices": opencl_devices(),
           "sample": "the generated smooth_scatter + smooth_gather kernels (project OpenCL-C bodies with the "
                     "engine's GMlib prologue/epilogue, one-row form) compiled for x86-64 by clang and run with "
                     "OpenMP over %d host threads, no CPU OpenCL runtime present; the same %d^3 mesh and initial "
                     "coordinates, 1 warm-up + %d timed Jacobi sweeps, median sweep" % (hk.threads, n, len(times))}
    return rec, c, len(times) + 1


def run_smoothing(args):
    """The scale workload: both paths from the same initial coordinates, each
    timed, each checked against the oracle's FP32 sweeps after every sweep
    it ran (the oracle run is also the CPU baseline)."""
    import numpy as np
    import gmlib_amd as G
    from tests import oracle_lib as O

    n = args.n
    t0 = time.time()
    crd, ref, tet = G.cube_mesh(n, 0.2, 2024, True)
    V, T = len(crd), len(tet)
    crd0 = np.ascontiguousarray(crd, np.float32)
    log("mesh %d^3: V=%d T=%d built in %.1fs" % (n, V, T, time.time() - t0))

    g = new_handle(args)
    L, h = g.L, g.h
    t0 = time.time()
    vi, ti = g.set_mesh(crd, ref, tet)
    par = g.new_parameters(16, body("smooth_params"))
    C.cast(par, C.POINTER(C.c_float))[0] = args.omega
    L.GmlUploadParameters(h)
    bary = g.solution(G.GmlTetrahedra, 1, G.GmlFlt4, "Bary")
    qal = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "qal")
    ks = g.compile(body("smooth_scatter"), "smooth_scatter", G.GmlTetrahedra,
                   [(vi, G.GmlReadMode, 0), (bary, G.GmlWriteMode, 0)])
    kg = g.compile(body("smooth_gather"), "smooth_gather", G.GmlVertices,
                   [(bary, G.GmlReadMode, 0), (vi, G.GmlReadMode | G.GmlWriteMode | G.GmlRefFlag, 0)])
    kq = g.compile(body("tet_meanratio"), "tet_meanratio", G.GmlTetrahedra,
                   [(vi, G.GmlReadMode, 0), (qal, G.GmlWriteMode, 0)])
    L.GmlxSync(h)
    setup_s = time.time() - t0
    log("upload + ball build + compile: %.1fs" % setup_s)

    if args.copy_calib:
        vdat = g.solution(G.GmlVertices, 1, G.GmlFlt4, "VerDat")
        kc = g.compile("VerDat = VerCrd;", "copy_calib", G.GmlVertices,
                       [(vi, G.GmlReadMode, 0), (vdat, G.GmlWriteMode, 0)])
        for _ in range(args.copy_calib):
            L.GmlLaunchKernel(h, kc)
        L.GmlxSync(h)
        log("copy_calib: %d launches of %d bytes read + %d bytes written" % (args.copy_calib, 16 * V, 16 * V))

    def restore():
        """the initial coordinates back into the vertex table"""
        if not L.GmlSetDataBlock(h, G.GmlVertices, 0, V - 1, G.ptr(crd0),
                                 C.c_void_p(crd0.ctypes.data + (V - 1) * 16), G.ptr(ref),
                                 C.c_void_p(ref.ctypes.data + (V - 1) * 4)):
            raise RuntimeError("GmlSetDataBlock failed")
        L.GmlxSync(h)

    # algorithmic bytes (each datum once) of the drop-in step, which moves
    # exactly the byte model of the workload: scatter 32T + 16V, gather
    # 32T + 40V, quality 20T + 16V, min 4T
    scat_bytes = 32 * T + 16 * V
    gath_bytes = 32 * T + 40 * V
    qual_bytes = 20 * T + 16 * V
    min_bytes = 4 * T
    model_step = args.sweeps * (scat_bytes + gath_bytes) + qual_bytes + min_bytes

    def step_generated():
        for _ in range(args.sweeps):
            if L.GmlLaunchKernel(h, ks) != 1 or L.GmlLaunchKernel(h, kg) != 1:
                raise RuntimeError("launch failed")
        if L.GmlLaunchKernel(h, kq) != 1:
            raise RuntimeError("launch failed")
        return g.reduce(qal, G.GmlMin)

    def step_builtin():
        if L.GmlxSmoothSweeps(h, bary, args.sweeps, args.omega) != 1 or L.GmlxTetQuality(h, qal) != 1:
            raise RuntimeError("built-in launch failed")
        return g.reduce(qal, G.GmlMin)

    def ktimes(path):
        if path == "generated":
            return [L.GmlGetKernelRunTime(h, k) for k in (ks, kg, kq)]
        cnt = C.c_longlong()
        return [L.GmlxGetBuiltinRunTime(h, w, C.byref(cnt))
                for w in (G.GmlxBuiltinSweep, G.GmlxBuiltinQuality)]

    L.GmlxSetBuiltinShape(h, G.GmlxBuiltinSweep, args.sweep_tile, 1)
    if args.quality_shape:
        qg, qp = (int(x) for x in args.quality_shape.split("/"))
        if not (L.GmlxSetBuiltinShape(h, G.GmlxBuiltinQuality, qg, 1) and L.GmlxSetBuiltinItems(h, G.GmlxBuiltinQuality, qp)):
            raise SystemExit("invalid --quality-shape %s" % args.quality_shape)
    paths, final = {}, {}
    nsteps = args.warmup + 1 + args.steps          # steps each path runs from the initial coordinates
    for path, step in (("generated", step_generated), ("builtin", step_builtin)):
        if path not in args.paths.split(","):
            continue
        restore()
        # the first step also builds the tile plans the kernels' first
        # launches need (element tiles, staging, the sweep plan): timed apart
        t0 = time.time()
        first_step = None
        for i in range(args.warmup):
            qmin = step()
            if i == 0:
                L.GmlxSync(h)
                first_step = time.time() - t0
        if path == "generated":
            # the launch-shape tuner of a generated kernel tries its shapes on
            # the first launches: finish it for the quality kernel too, which a
            # step launches once, so that the timed steps are steady-state
            # (quality launches do not move the coordinates)
            while L.GmlxGetKernelRows(h, kq) == 0:
                if L.GmlLaunchKernel(h, kq) != 1:
                    raise RuntimeError("launch failed")
        qmin = step()
        sg = ktimes(path)
        L.GmlxSync(h)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            qmin = step()
        L.GmlxSync(h)
        dt = time.perf_counter() - t0
        ek = [a - b for a, b in zip(ktimes(path), sg)]
        nl = args.steps * args.sweeps
        forms = {}
        if path == "generated":
            per = {"scatter": (scat_bytes, ek[0] / nl), "gather": (gath_bytes, ek[1] / nl),
                   "quality": (qual_bytes, ek[2] / args.steps)}
            # a kernel in the element-tile form reads its tile plan (index
            # words per tet, tile offsets, tile vertex lists) and each tile's
            # vertices once instead of its tet rows and the scattered vertex
            # gathers: its own compulsory bytes replace the model's 16T + 16V
            pi = (C.c_longlong * 6)()
            have_plan = L.GmlxGetTilePlanInfo(h, G.GmlTetrahedra, pi) == 1
            own = {}
            for name, k, outb in (("scatter", ks, 16 * T), ("gather", kg, None), ("quality", kq, 4 * T)):
                forms[name] = L.GmlxGetKernelForm(h, k)
                if forms[name] == 3 and have_plan and outb is not None:
                    own[name] = 4 * pi[4] * T + 4 * (pi[1] + 1) + 20 * pi[2] + outb
            own_step = (args.sweeps * (own.get("scatter", scat_bytes) + own.get("gather", gath_bytes))
                        + own.get("quality", qual_bytes) + min_bytes)
        else:
            info = (C.c_longlong * 9)()
            if not L.GmlxGetSweepPlanInfo(h, info):
                raise RuntimeError("no sweep plan")
            log("sweep plan: mode %d, tile %d, %d tiles, tile vertices / V = %.3f, 1-ring slots %d, %.2f GB"
                % (info[8], info[0], info[1], info[2] / V, info[7], info[6] / 1e9))
            # operator form: 1-ring slots (2 B; the degree is their count sum
            # / 3) + ref (4 B) + old and new coordinates (16 B each); exact
            # fused form: tet->vertex rows as the plan's 8 B + ball slots 2 B
            # x 4T + 40 V
            sweep_alg = 2 * info[7] + 36 * V if info[8] == 2 else 16 * T + 40 * V
            per = {"sweep": (sweep_alg, ek[0] / nl), "quality": (qual_bytes, ek[1] / args.steps)}
            own_step = args.sweeps * sweep_alg + qual_bytes + min_bytes
        kern = {k: {"avg_us": round(v[1] * 1e6, 2), "alg_GBs": round(v[0] / v[1] / 1e9, 1),
                    "frac": round(v[0] / v[1] / 1e9 / HBM_PEAK_GBS, 4), "bytes_per_launch": v[0]}
                for k, v in per.items()}
        if path == "generated":
            for k in kern:
                kern[k]["form"] = {1: "one row", 2: "paired", 3: "element tiles"}.get(forms[k], str(forms[k]))
                if k in own:
                    t = per[k][1]
                    kern[k].update({"own_bytes_per_launch": own[k], "own_GBs": round(own[k] / t / 1e9, 1),
                                    "own_frac": round(own[k] / t / 1e9 / HBM_PEAK_GBS, 4)})
        ms = dt / args.steps * 1e3
        rec = {"ms_per_step": round(ms, 4), "first_step_s": round(first_step, 3) if first_step is not None else None,
               "own_bytes_per_step": own_step,
               "GBs": round(own_step / (dt / args.steps) / 1e9, 2),
               "model_equivalent_GBs": round(model_step / (dt / args.steps) / 1e9, 2),
               "kernels": kern, "min_quality": qmin}
        # the final state: coordinates after nsteps steps, quality of the last
        c, _ = g.vertices(V)
        q = g.download_raw(qal, np.float32, (T,))
        final[path] = (c[:, :3].copy(), q, qmin)
        paths[path] = rec
        log("%s: step %.3f ms, %.1f GB/s of its own bytes (%.1f GB/s model-equivalent); kernels %s; min quality %.6f"
            % (path, ms, rec["GBs"], rec["model_equivalent_GBs"], kern, qmin))
    shapes = []
    for k in (ks, kg, kq):
        a, b = C.c_int(), C.c_int()
        L.GmlxGetKernelShape(h, k, C.byref(a), C.byref(b))
        shapes.append((a.value, b.value))
    log("generated kernel shapes (group, remap): %s" % shapes)
    if "generated" in paths and "builtin" in paths:
        paths["builtin"]["speedup_vs_dropin"] = round(paths["generated"]["ms_per_step"] / paths["builtin"]["ms_per_step"], 2)
    head = "generated" if "generated" in paths else "builtin"
    P = paths[head]
    dom = max(P["kernels"], key=lambda k: P["kernels"][k]["avg_us"] * (args.sweeps if k != "quality" else 1))
    D = P["kernels"][dom]
    # HBM traffic per launch of the dominant kernel, as measured by the
    # committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs on this workload
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_256cube.json")
    if n == 256 and os.path.exists(tf):
        key = dom if dom != "quality" else ("quality_builtin" if head == "builtin" else "quality_generated")
        traffic = json.load(open(tf)).get(key)
    ball_w = 32

    # the oracle from the same initial coordinates, as many FP32 sweeps as
    # each path ran (the parity check); the CPU baseline is the same
    # generated kernels built for the host cores (oracle/host_kernels.py)
    S = nsteps * args.sweeps
    deg = np.empty(V, np.int32)
    ball = np.empty((V, ball_w), np.int32)
    if not L.GmlxGetUplinkTable(h, G.GmlVertices, G.GmlTetrahedra, G.ptr(deg), G.ptr(ball), ball_w):
        raise RuntimeError("uplink download failed")
    gpu_gather_src = L.GmlxGetKernelSource(h, kg).decode()
    g.close()
    cpu, host_c, KH = host_baseline(crd0, ref, tet, deg, ball, gpu_gather_src, args.omega, scat_bytes + gath_bytes, n)
    OL = O.lib()
    c = np.array(crd0, copy=True)
    b4 = np.empty((T, 4), np.float32)
    cpu["equal_to_oracle"] = {"sweeps": KH, "equal": None}
    t0 = time.perf_counter()
    for i in range(S):
        OL.orc_smooth_sweep_f32(V, T, O.p(tet), O.p(ref), O.p(deg), O.p(ball), ball_w, O.p(c), O.p(b4), args.omega)
        if i + 1 == KH:
            # the host-built generated kernels against the oracle after as many sweeps
            cpu["equal_to_oracle"]["equal"] = bool(np.array_equal(host_c[:, :3], c[:, :3]))
    ct = time.perf_counter() - t0
    cpu["oracle_c"] = {"value": round(S * (scat_bytes + gath_bytes) / ct / 1e9, 2), "unit": "GB/s",
                       "cores": OL.orc_num_threads(),
                       "sample": "%d Jacobi sweeps (scatter + gather), the oracle's FP32 C + OpenMP loops, "
                                 "%.2f s" % (S, ct)}
    log("cpu baseline: %s" % cpu)
    parity = {"oracle": "orc_smooth_sweep_f32, %d sweeps from the same initial coordinates" % S,
              "coord_tolerance": 1e-5, "quality_rel_tolerance": 1e-5}
    ok = True
    for path, (cg, qg, qmin) in final.items():
        dx = float(np.abs(cg - c[:, :3]).max())
        # min quality: the GPU reduction against the same FP32 array
        min_exact = qmin == float(qg.min())
        # quality against FP64 on the GPU's own final coordinates
        q64 = O.tet_quality64(tet, np.hstack([cg, np.zeros((V, 1), np.float32)]))
        big = np.abs(q64) >= 0.05
        qerr = float(np.max(np.abs(qg[big] - q64[big]) / np.abs(q64[big])))
        parity[path] = {"max_abs_dx_vs_oracle": dx, "min_quality_bitexact": bool(min_exact),
                        "quality_max_rel_err_vs_fp64": qerr}
        ok &= dx <= 1e-5 and min_exact and qerr <= 1e-5
    parity["ok"] = bool(ok)
    log("parity: %s" % parity)

    return {
        "metric": "HBM GB/s (% peak) per kernel; served mesh-jobs/s + p99 vs plain path",
        "value": P["GBs"],
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": P["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded jittered Kuhn cube mesh)",
        "config": {"workload": "Laplacian node smoothing, %d^3-cell Kuhn cube (%d vertices, %d tets), "
                               "%d Jacobi sweeps + tet quality + min reduction, drop-in C API path "
                               "(GmlCompileKernel / GmlLaunchKernel / GmlReduceVector)" % (n, V, T, args.sweeps),
                   "vertices": V, "tets": T, "sweeps": args.sweeps, "path": head,
                   "bytes_per_step": P["own_bytes_per_step"],
                   "setup_s": {"upload_ball_compile": round(setup_s, 3),
                               "first_step": {k: v["first_step_s"] for k, v in paths.items()}}},
        "roofline": {"bound": "hbm", "achieved": D["alg_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": D["frac"], "traffic": traffic, "kernel": "%s (%s path)" % (dom, head),
                     "bytes_per_launch": D["bytes_per_launch"], "avg_us": D["avg_us"]},
        "paths": paths,
        "parity": parity,
        "cpu_baseline": cpu,
    }


def run_small(args):
    """The CPU-runnable configuration: tet quality, ball minimum and
    GmlReduceVector min / mean on a 20^3-cell cube (48,000 tets), through the
    drop-in generated kernels and the hand-written ones, in microseconds and
    GB/s, checked against the oracle."""
    import math
    import numpy as np
    import gmlib_amd as G
    from tests import oracle_lib as O

    n = 20
    crd, ref, tet = G.cube_mesh(n, 0.2, 2020, True)
    V, T = len(crd), len(tet)
    qb = 16 * T + 16 * V + 4 * T
    bb = 16 * T + 4 * V + 4 * T + 4 * V
    rb = 4 * T
    g = new_handle(args)
    L, h = g.L, g.h
    vi, ti = g.set_mesh(crd, ref, tet)
    q = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "qal")
    vq = g.solution(G.GmlVertices, 1, G.GmlFlt, "vq")
    kq = g.compile(body("tet_meanratio"), "tet_meanratio", G.GmlTetrahedra, [(vi, G.GmlReadMode, 0), (q, G.GmlWriteMode, 0)])
    kb = g.compile(body("ball_min"), "ball_min", G.GmlVertices, [(q, G.GmlReadMode, 0), (vq, G.GmlWriteMode, 0)])
    red = C.c_double()

    def step(path):
        if path == "generated":
            if L.GmlLaunchKernel(h, kq) != 1 or L.GmlLaunchKernel(h, kb) != 1:
                raise RuntimeError("launch failed")
        elif L.GmlxTetQuality(h, q) != 1 or L.GmlxBallMin(h, q, vq) != 1:
            raise RuntimeError("built-in launch failed")
        L.GmlReduceVector(h, q, G.GmlMin, C.byref(red))
        mn = red.value
        L.GmlReduceVector(h, q, G.GmlSum, C.byref(red))
        return mn, red.value / T

    def times(path):
        cnt = C.c_longlong()
        if path == "generated":
            k = [L.GmlGetKernelRunTime(h, kq), L.GmlGetKernelRunTime(h, kb)]
        else:
            k = [L.GmlxGetBuiltinRunTime(h, G.GmlxBuiltinQuality, C.byref(cnt)),
                 L.GmlxGetBuiltinRunTime(h, G.GmlxBuiltinBallMin, C.byref(cnt))]
        return k + [L.GmlGetReduceRunTime(h, G.GmlMin)]

    steps = max(args.steps, 50)
    res = {}
    for path in ("generated", "builtin"):
        for _ in range(20):                       # the generated kernels' shape tuners settle
            step(path)
        t0 = times(path)
        w0 = time.perf_counter()
        for _ in range(steps):
            mn, mean = step(path)
        wall = (time.perf_counter() - w0) / steps
        dk = [(a - b) / steps for a, b in zip(times(path), t0)]
        qg = g.download_raw(q, np.float32, (T,))
        vg = g.download_raw(vq, np.float32, (V,))
        res[path] = {"step_us": round(wall * 1e6, 2),
                     "quality": {"avg_us": round(dk[0] * 1e6, 2), "GBs": round(qb / dk[0] / 1e9, 1)},
                     "ball_min": {"avg_us": round(dk[1] * 1e6, 2), "GBs": round(bb / dk[1] / 1e9, 1)},
                     "reduce_min": {"avg_us": round(dk[2] * 1e6, 2), "GBs": round(rb / dk[2] / 1e9, 1)},
                     "min_quality": mn, "mean_quality": mean, "_q": qg, "_v": vg}
    g.close()
    # parity: quality vs FP64, ball minimum and min bit-exact on the GPU's own
    # FP32 qualities, mean vs the exact sum
    q64 = O.tet_quality64(tet, crd)
    deg, ball, _ = O.vertex_ball(tet, V, 32)
    par = {}
    ok = True
    for path, r in res.items():
        qg, vg = r.pop("_q"), r.pop("_v")
        qerr = float(np.max(np.abs(qg - q64) / np.maximum(np.abs(q64), 1e-3)))
        ball_exact = bool(np.array_equal(vg, O.ball_min(deg, ball, qg)))
        min_exact = r["min_quality"] == float(qg.min())
        mean_err = abs(r["mean_quality"] - math.fsum(q64) / T) / abs(math.fsum(q64) / T)
        par[path] = {"quality_max_rel_err_vs_fp64": qerr, "ball_min_bitexact": ball_exact,
                     "min_bitexact": bool(min_exact), "mean_rel_err_vs_fp64": mean_err}
        ok &= qerr <= 1e-5 and ball_exact and min_exact and mean_err <= 1e-5
    par["ok"] = bool(ok)
    # CPU baseline: the same generated quality and ball-minimum kernels built
    # for the host cores (oracle/host_kernels.py), the reductions by the
    # oracle's C loops; the oracle's own quality and ball minimum beside it
    from oracle import host_kernels as HK
    OL = O.lib()
    src = HK.quality_sources(G, os.path.join(ROOT, "gmlib_amd", "kernels"))
    hk = HK.HostKernels(HK.build(src, tag="quality"), src)
    hq = HK.Quality(hk, crd, tet, deg, ball)

    def host_step():
        hq.step()
        O.reduce(G.GmlMin, hq.qal)
        O.reduce(G.GmlSum, hq.qal)

    def oracle_step():
        q32 = O.tet_quality32(tet, crd)
        O.ball_min(deg, ball, q32)
        O.reduce(G.GmlMin, q32)
        O.reduce(G.GmlSum, q32)

    def per_step(fn):
        fn()
        t0, k = time.perf_counter(), 0
        while k < 5 or time.perf_counter() - t0 < min(args.cpu_seconds, 2.0):
            fn()
            k += 1
        return (time.perf_counter() - t0) / k, k

    ct, k = per_step(host_step)
    q32 = O.tet_quality32(tet, crd)
    host_equal = bool(np.array_equal(hq.qal, q32) and np.array_equal(hq.vq, O.ball_min(deg, ball, q32)))
    ot, ko = per_step(oracle_step)
    out = {"workload": "tet quality + ball min + GmlReduceVector min / mean, 20^3-cell Kuhn cube (%d vertices, "
                       "%d tets)" % (V, T),
           "vertices": V, "tets": T, "bytes": {"quality": qb, "ball_min": bb, "reduce": rb},
           "paths": res, "parity": par,
           "cpu_baseline": {"value": round((qb + bb + 2 * rb) / ct / 1e9, 2), "unit": "GB/s", "step_us": round(ct * 1e6, 1),
                            "cores": hk.threads, "kind": "port", "equal_to_oracle": host_equal,
                            "sample": "%d steps: the generated tet_meanratio and ball_min kernels (project OpenCL-C "
                                      "bodies with the engine's GMlib prologue/epilogue) compiled for x86-64 by clang "
                                      "and run with OpenMP, min and sum by the oracle's C loops" % k,
                            "oracle_c": {"value": round((qb + bb + 2 * rb) / ot / 1e9, 2), "step_us": round(ot * 1e6, 1),
                                         "cores": OL.orc_num_threads(),
                                         "sample": "%d oracle steps (FP32 quality, ball min, min and sum), "
                                                   "C + OpenMP" % ko}}}
    log("config 1 (20^3): %s" % out)
    return out


def run_cg(args):
    import numpy as np
    import gmlib_amd as G
    from tests import oracle_lib as O

    n = args.cg_n
    t0 = time.time()
    crd, ref, tet = G.cube_mesh(n, 0.2, 2024, True)
    L0 = G.lib()
    rp = np.empty(len(crd) + 1, np.int32)
    nb = L0.GmlxEdgeLaplacian3(len(crd), len(tet), G.ptr(tet), 2024, 0.01, G.ptr(rp), None, None, 0)
    col = np.empty(nb, np.int32)
    val = np.empty(nb * 9, np.float32)
    if L0.GmlxEdgeLaplacian3(len(crd), len(tet), G.ptr(tet), 2024, 0.01, G.ptr(rp), G.ptr(col), G.ptr(val), nb) != nb:
        raise RuntimeError("GmlxEdgeLaplacian3 failed")
    del crd, ref, tet
    nv, nb = len(rp) - 1, len(col)
    rhs = np.random.default_rng(7).normal(size=nv * 3).astype(np.float32)
    log("edge Laplacian %d^3: %d block rows, %d 3x3 blocks (%.1f per row), built in %.1fs"
        % (n, nv, nb, nb / nv, time.time() - t0))
    spmv_bytes = 40 * nb + 28 * nv

    g = new_handle(args)
    L, h = g.L, g.h
    m = L.GmlNewMatrix(h, nv, nb, 3, G.ptr(val), G.ptr(col), G.ptr(rp), G.GmlFlt)
    vx = L.GmlNewVector(h, nv, 3, G.ptr(rhs), G.GmlFlt)
    vy = L.GmlNewVector(h, nv, 3, None, G.GmlFlt)
    vs = L.GmlNewVector(h, nv, 3, None, G.GmlDbl)
    if min(m, vx, vy, vs) <= 0:
        raise RuntimeError("matrix/vector creation failed")
    # the product's own bytes: a block whose column the kernel reads as a
    # 16-bit offset from its row (GmlxGetMatrixInfo) costs 38 B instead of 40
    minfo = (C.c_longlong * 6)()
    if L.GmlxGetMatrixInfo(h, m, minfo) != 1:
        raise RuntimeError("GmlxGetMatrixInfo failed")
    spmv_own = spmv_bytes - 2 * minfo[5]
    log("matrix: %d SELL-64 slabs, %d of them with 16-bit columns; %d of %d blocks (%.1f%%) read through 16-bit "
        "columns" % (minfo[3], minfo[4], minfo[5], nb, 100. * minfo[5] / nb))
    if args.copy_calib:
        one = C.c_double(1.0)
        for _ in range(args.copy_calib):
            L.GmlScaleVec(h, vy, C.byref(one))
        L.GmlxSync(h)
        log("copy_calib: %d GmlScaleVec launches (k_scale) of %d bytes read + %d bytes written"
            % (args.copy_calib, 12 * nv, 12 * nv))

    # SpMV: the step.  Launch shapes (work-group size, XCD remap) are swept
    # first on a few launches each; the best one is timed.
    cnt = C.c_longlong()

    def time_spmv(k):
        L.GmlxSync(h)
        k0 = L.GmlxGetBuiltinRunTime(h, G.GmlxBuiltinSpMV, C.byref(cnt))
        c0 = cnt.value
        t0 = time.perf_counter()
        for _ in range(k):
            if L.GmlMultMatVec(h, m, vy, vx) != 1:
                raise RuntimeError("GmlMultMatVec failed")
        L.GmlxSync(h)
        dt = (time.perf_counter() - t0) / k
        k1 = L.GmlxGetBuiltinRunTime(h, G.GmlxBuiltinSpMV, C.byref(cnt))
        return dt, (k1 - k0) / (cnt.value - c0)

    shapes = {}
    for grp in (64, 128, 256):
        for rm in (0, 1):
            L.GmlxSetBuiltinShape(h, G.GmlxBuiltinSpMV, grp, rm)
            time_spmv(2)
            shapes["%d/%d" % (grp, rm)] = time_spmv(5)[1]
    best = min(shapes, key=shapes.get)
    grp, rm = (int(v) for v in best.split("/"))
    log("SpMV shapes (group/remap: kernel us): %s -> %s"
        % ({k: round(v * 1e6, 1) for k, v in shapes.items()}, best))
    for w in (G.GmlxBuiltinSpMV, G.GmlxBuiltinCgSpMV):
        L.GmlxSetBuiltinShape(h, w, grp, rm)
    for _ in range(args.warmup):
        L.GmlMultMatVec(h, m, vy, vx)
    dt, kavg = time_spmv(args.steps)
    y = np.empty(nv * 3, np.float32)
    L.GmlxGetVector(h, vy, G.ptr(y))
    y64 = O.bsr_spmv64(rp, col, val, rhs.astype(np.float64), 3)
    spmv_err = float(np.abs(y - y64).max() / np.abs(y64).max())
    gbs = spmv_bytes / dt / 1e9
    kgbs = spmv_bytes / kavg / 1e9
    log("SpMV: %.1f us/step (kernel %.1f us), %.1f GB/s step, %.1f GB/s kernel (%.1f%% of peak), "
        "max err / max|y| vs FP64 oracle %.2e" % (dt * 1e6, kavg * 1e6, gbs, kgbs, kgbs / HBM_PEAK_GBS * 100, spmv_err))

    # the two CG forms, each checked with the oracle's FP64 residual
    # per CG iteration, every datum once: SpMV kernel (q = A p, p.q) 40 B per
    # block + 28 B per row (row length, p read, q written); update kernel 108 B
    # per row (x, p, r, q, 3x3 Minv read; x, r, z written) + direction update
    # 36 B per row (z, p read; p written)
    cg_spmv_bytes = 40 * nb + 28 * nv
    cg_upd_bytes = 144 * nv
    bn = float(np.linalg.norm(rhs.astype(np.float64)))
    solves = {}
    for name, tol, refine in (("fp32", 1e-5, 0), ("fp32_fp64_refined", 1e-10, 1)) if args.cg_solve else ():
        rr, it = C.c_double(), C.c_int()
        ca, cb = C.c_longlong(), C.c_longlong()
        s0 = L.GmlxGetBuiltinRunTime(h, G.GmlxBuiltinCgSpMV, C.byref(ca)); n0 = ca.value
        u0 = L.GmlxGetBuiltinRunTime(h, G.GmlxBuiltinCgUpdate, C.byref(cb))
        t0 = time.perf_counter()
        rc = L.GmlxSolveCG(h, m, vx, vs, tol, 20000, refine, C.byref(rr), C.byref(it))
        ts = time.perf_counter() - t0
        if rc != 1:
            raise RuntimeError("GmlxSolveCG(%s) returned %d" % (name, rc))
        s1 = L.GmlxGetBuiltinRunTime(h, G.GmlxBuiltinCgSpMV, C.byref(ca))
        u1 = L.GmlxGetBuiltinRunTime(h, G.GmlxBuiltinCgUpdate, C.byref(cb))
        nl = max(1, ca.value - n0)
        x = np.empty(nv * 3, np.float64)
        L.GmlxGetVector(h, vs, G.ptr(x))
        res = float(np.linalg.norm(rhs.astype(np.float64) - O.bsr_spmv64(rp, col, val, x, 3)) / bn)
        solves[name] = {"tol": tol, "iterations": it.value, "seconds": round(ts, 4),
                        "ms_per_iteration": round(ts / max(1, it.value) * 1e3, 4),
                        "relres_gpu_fp64": rr.value, "relres_oracle_fp64": res,
                        "cg_spmv_us": round((s1 - s0) / nl * 1e6, 2),
                        "cg_spmv_GBs": round(cg_spmv_bytes * nl / (s1 - s0) / 1e9, 1),
                        "cg_update_us": round((u1 - u0) / nl * 1e6, 2),
                        "cg_update_GBs": round(cg_upd_bytes * nl / (u1 - u0) / 1e9, 1),
                        "kernel_share": round((s1 - s0 + u1 - u0) / ts, 3)}
        log("CG %s: %s" % (name, solves[name]))
        if res > max(2 * tol, 1e-9) * (10 if not refine else 1):
            raise RuntimeError("CG %s residual %.3e above its bound" % (name, res))

    # plain streaming rate of this GPU: GmlScaleVec (read + write) over a 2 GiB vector
    # (skipped in counter-calibration runs, whose GmlScaleVec launches must all be the small one)
    big = L.GmlNewVector(h, 1 << 27, 4, None, G.GmlFlt) if not args.copy_calib and args.cg_solve else 0
    probe = None
    if big > 0:
        one = C.c_double(1.0)
        for _ in range(3):
            L.GmlScaleVec(h, big, C.byref(one))
        L.GmlxSync(h)
        t0 = time.perf_counter()
        for _ in range(10):
            L.GmlScaleVec(h, big, C.byref(one))
        L.GmlxSync(h)
        probe = round(2 * (1 << 29) * 4 * 10 / (time.perf_counter() - t0) / 1e9, 1)
        L.GmlxFreeVector(h, big)
        log("streaming probe (GmlScaleVec, 2 GiB read + 2 GiB written): %.1f GB/s" % probe)

    cpu = None
    if args.cpu_seconds > 0:
        OL = O.lib()
        y32 = np.empty(nv * 3, np.float32)
        OL.orc_bsr_spmv_f32(nv, 3, O.p(rp), O.p(col), O.p(val), O.p(rhs), O.p(y32))
        t0 = time.perf_counter()
        ns = 0
        while ns < 3 or time.perf_counter() - t0 < args.cpu_seconds:
            OL.orc_bsr_spmv_f32(nv, 3, O.p(rp), O.p(col), O.p(val), O.p(rhs), O.p(y32))
            ns += 1
        ct = time.perf_counter() - t0
        cpu = {"value": round(ns * spmv_bytes / ct / 1e9, 2), "unit": "GB/s",
               "cores": OL.orc_num_threads(), "kind": "port",
               "sample": "%d block-CSR SpMVs of the same matrix, oracle FP32 C + OpenMP, %.2f s" % (ns, ct)}
        log("cpu baseline: %s" % cpu)

    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_cg%d.json" % n)
    if os.path.exists(tf):
        traffic = json.load(open(tf)).get("spmv")
    out = {
        "metric": "HBM GB/s (% peak) per kernel; served mesh-jobs/s + p99 vs plain path",
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SPD edge Laplacian of a jittered Kuhn cube mesh, N(0,1) right-hand side)",
        "config": {"workload": "3x3 block SpMV + block-Jacobi CG, edge Laplacian of a %d^3-cell cube "
                               "(%d block rows, %d blocks)" % (n, nv, nb),
                   "block_rows": nv, "blocks": nb, "bytes_per_step": spmv_bytes, "spmv_max_rel_err": spmv_err,
                   "spmv_shape": best, "spmv_shapes_us": {k: round(v * 1e6, 2) for k, v in shapes.items()}},
        "roofline": {"bound": "hbm", "achieved": round(kgbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(kgbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_spmv<3,float,float> (GmlMultMatVec)",
                     "bytes_per_launch": spmv_bytes, "avg_us": round(kavg * 1e6, 2),
                     "own_bytes_per_launch": spmv_own, "own_frac": round(spmv_own / kavg / 1e9 / HBM_PEAK_GBS, 4),
                     "blocks_16bit_columns": minfo[5]},
        "cg": solves,
        "stream_probe_GBs": probe,
        "cpu_baseline": cpu,
    }
    g.close()
    return out


def serve_stream(args, seed=2025):
    """(pool of (n, seed) meshes, job list of (type, pool index, param), arrival times)."""
    import numpy as np
    import gmlib_amd as G
    rng = np.random.default_rng(seed)
    tets = np.exp(rng.uniform(np.log(1e4), np.log(2e7), args.pool))
    pool = [(max(2, int(round((t / 6) ** (1 / 3)))), 9000 + i) for i, t in enumerate(tets)]
    kinds = rng.choice([G.GmlxJobQuality, G.GmlxJobSmooth, G.GmlxJobCG], size=args.jobs, p=[0.5, 0.3, 0.2])
    which = rng.integers(len(pool), size=args.jobs)
    param = {G.GmlxJobQuality: 0.0, G.GmlxJobSmooth: 20.0, G.GmlxJobCG: 1e-8}
    jobs = [(int(k), int(w), param[int(k)]) for k, w in zip(kinds, which)]
    arrivals = np.cumsum(rng.exponential(1.0 / args.qps, size=args.jobs))
    return pool, jobs, arrivals


def serve_run(L, srv, keys, jobs, arrivals=None):
    """Submit jobs on registered meshes (at their arrival times, or all at
    once) and wait for all: list of (rc, res[4], tim[4])."""
    import threading
    ids = [None] * len(jobs)
    t0 = time.perf_counter()

    def submitter():
        for i, (t, w, p) in enumerate(jobs):
            if arrivals is not None:
                d = t0 + arrivals[i] - time.perf_counter()
                if d > 0:
                    time.sleep(d)
            ids[i] = L.GmlxServeSubmitMesh(srv, t, keys[w], p)

    th = threading.Thread(target=submitter)
    th.start()
    out = []
    for i in range(len(jobs)):
        while ids[i] is None:
            time.sleep(0.0005)
        res, tim = (C.c_double * 4)(), (C.c_double * 4)()
        rc = L.GmlxServeWait(srv, ids[i], res, tim)
        out.append((rc, tuple(res), tuple(tim)))
    th.join()
    return out


def latency_stats(out):
    import numpy as np
    lat = np.array([o[2][2] - o[2][0] for o in out])
    svc = np.array([o[2][2] - o[2][1] for o in out])
    span = max(o[2][2] for o in out) - min(o[2][0] for o in out)
    return {"jobs_per_s": round(len(out) / span, 2), "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 2),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 2), "mean_ms": round(float(lat.mean()) * 1e3, 2),
            "service_p50_ms": round(float(np.percentile(svc, 50)) * 1e3, 2),
            "service_p99_ms": round(float(np.percentile(svc, 99)) * 1e3, 2)}


def register_pool(L, srv, meshes):
    keys = []
    for crd, ref, tet in meshes:
        k = L.GmlxServeAddMesh(srv, len(crd), len(tet), G_ptr(crd), G_ptr(ref), G_ptr(tet))
        if k <= 0:
            raise RuntimeError("GmlxServeAddMesh failed")
        keys.append(k)
    return keys


def G_ptr(a):
    return C.c_void_p(a.ctypes.data)


def run_serve(args):
    """Served stream against the plain per-job path.  Both servers get the
    same host meshes, generated before either runs and registered with
    GmlxServeAddMesh, and the plain server is warmed so that the host-side CG
    systems of its jobs exist: a plain job is GmlInit, upload, links and
    plans built on the device, run, reduce, GmlStop, one at a time."""
    import math
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    import gmlib_amd as G
    from tests import oracle_lib as O
    L = G.lib()
    pool, jobs, arrivals = serve_stream(args)
    tets = [6 * n ** 3 for n, _ in pool]
    t0 = time.time()
    with ThreadPoolExecutor(16) as ex:
        meshes = list(ex.map(lambda ps: G.cube_mesh(ps[0], 0.2, ps[1]), pool))
    log("serve: pool of %d meshes, %d to %d tets (%.1fM total), generated in %.1fs; %d jobs at %.1f jobs/s; "
        "%d slots, budget %.0f GB" % (len(pool), min(tets), max(tets), sum(tets) / 1e6, time.time() - t0, len(jobs),
                                       args.qps, args.slots, args.budget_gb))
    srv = L.GmlxServeStart(0, args.budget_gb, args.slots, 1)
    if not srv:
        raise RuntimeError("GmlxServeStart failed")
    keys = register_pool(L, srv, meshes)
    # warm-up: every mesh loaded, its smoothing plan and CG system built
    t0 = time.time()
    warm = [(k, w, p) for w in range(len(pool)) for k, p in ((G.GmlxJobSmooth, 20.0), (G.GmlxJobCG, 1e-8))]
    wout = serve_run(L, srv, keys, warm)
    if any(o[0] != 1 for o in wout):
        raise RuntimeError("warm-up jobs failed: %s" % [o[0] for o in wout])
    log("warm-up (%d jobs: load, plans, CG systems): %.1fs" % (len(warm), time.time() - t0))
    # the timed open-loop stream at the fixed rate
    out = serve_run(L, srv, keys, jobs, arrivals)
    if any(o[0] != 1 for o in out):
        raise RuntimeError("served jobs failed: %s" % sorted(set(o[0] for o in out)))
    st = latency_stats(out)
    by_kind = {}
    for name, k in (("quality", G.GmlxJobQuality), ("smooth", G.GmlxJobSmooth), ("cg", G.GmlxJobCG)):
        sel = [o for o, j in zip(out, jobs) if j[0] == k]
        if sel:
            lat = np.array([o[2][2] - o[2][0] for o in sel])
            by_kind[name] = {"jobs": len(sel), "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 2),
                             "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 2)}
    info = (C.c_double * 8)()
    L.GmlxServeInfo(srv, info)
    log("served at %.1f jobs/s offered: %s; by kind %s; resident %d meshes, %.2f GB, %d loads, %d evictions, "
        "up to %d jobs at once" % (args.qps, st, by_kind, info[0], info[1] / 1e9, info[2], info[3], info[5]))
    # saturated: the same jobs submitted at once
    sat = latency_stats(serve_run(L, srv, keys, jobs))
    log("served, saturated (all %d jobs at once): %s" % (len(jobs), sat))
    L.GmlxServeStop(srv)

    # plain per-job path on the first jobs of the stream, results compared
    npl = min(len(jobs), 40)
    plain = L.GmlxServeStart(0, args.budget_gb, 1, 0)
    pkeys = register_pool(L, plain, meshes)
    cg_meshes = sorted(set(w for t, w, _ in jobs[:npl] if t == G.GmlxJobCG))
    t0 = time.time()
    if any(o[0] != 1 for o in serve_run(L, plain, pkeys, [(G.GmlxJobCG, w, 1e-8) for w in cg_meshes])):
        raise RuntimeError("plain warm-up failed")
    log("plain warm-up (host CG systems of %d meshes): %.1fs" % (len(cg_meshes), time.time() - t0))
    pout = serve_run(L, plain, pkeys, jobs[:npl])
    L.GmlxServeStop(plain)
    if any(o[0] != 1 for o in pout):
        raise RuntimeError("plain jobs failed")
    pst = latency_stats(pout)
    same = all(a[1][:2] == b[1][:2] for a, b in zip(pout, out[:npl]))
    log("plain per-job path (%d jobs, closed loop): %s; results equal to the served ones: %s" % (npl, pst, same))

    # served results against the oracle, on the jobs of meshes up to 2M tets:
    # quality (min bit-exact is not expected across FP32 and FP64 -- min and
    # mean within 1e-5 relative of the FP64 quality), smoothing (20 FP32
    # oracle sweeps, then FP64 quality of those coordinates), CG (the GPU's
    # FP64 relative residual at or below the tolerance)
    checks = {"quality": [], "smooth": [], "cg": []}
    cache = {}
    t0 = time.time()
    for (t, w, p), o in zip(jobs, out):
        if tets[w] > 2e6:
            continue
        crd, ref, tet = meshes[w]
        if t == G.GmlxJobCG:
            checks["cg"].append(o[1][0] <= p)
            continue
        key = (t, w)
        if key not in cache:
            if t == G.GmlxJobQuality:
                q = O.tet_quality64(tet, crd)
            else:
                deg, ball, _ = O.vertex_ball(tet, len(crd), 32)
                c = O.smooth32(tet, ref, deg, ball, crd, int(p), 0.2)
                q = O.tet_quality64(tet, c)
            cache[key] = (float(q.min()), math.fsum(q) / len(q))
        mn, mean = cache[key]
        name = "quality" if t == G.GmlxJobQuality else "smooth"
        checks[name].append((abs(o[1][0] - mn) / abs(mn), abs(o[1][1] - mean) / abs(mean)))

    def worst(v, i):
        return max(e[i] for e in v) if v else None
    # a served smoothing runs the operator-form sweep, equal to the oracle's
    # two-kernel sweep up to FP32 rounding of the coordinates; the minimum
    # quality moves by about that rounding over the edge length, so its bound
    # is 1e-4 where the mean's (errors average out) stays 1e-5
    oracle = {"quality_jobs": len(checks["quality"]),
              "quality_min_max_rel_err": worst(checks["quality"], 0),
              "quality_mean_max_rel_err": worst(checks["quality"], 1),
              "smooth_jobs": len(checks["smooth"]),
              "smooth_min_max_rel_err": worst(checks["smooth"], 0),
              "smooth_mean_max_rel_err": worst(checks["smooth"], 1),
              "cg_jobs": len(checks["cg"]), "cg_all_at_tolerance": all(checks["cg"]),
              "tolerances": {"quality": 1e-5, "smooth_min": 1e-4, "smooth_mean": 1e-5},
              "seconds": round(time.time() - t0, 1)}
    oracle["ok"] = bool(all(a <= 1e-5 and b <= 1e-5 for a, b in checks["quality"])
                        and all(a <= 1e-4 and b <= 1e-5 for a, b in checks["smooth"]) and all(checks["cg"]))
    log("served results vs the oracle: %s" % oracle)
    return {
        "metric": "HBM GB/s (% peak) per kernel; served mesh-jobs/s + p99 vs plain path",
        "value": st["jobs_per_s"],
        "unit": "jobs/s",
        "n_gpus": 1,
        "steps": len(jobs),
        "warmup": len(warm),
        "ms_per_step": round(1e3 / st["jobs_per_s"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded cube meshes, seeded Poisson job stream)",
        "config": {"workload": "served stream of mixed mesh jobs (50%% quality, 30%% 20-sweep smoothing, 20%% CG to "
                               "1e-8), %d-mesh pool of 10k-20M tets, %.1f jobs/s offered, %d slots, %.0f GB budget"
                               % (len(pool), args.qps, args.slots, args.budget_gb),
                   "qps": args.qps, "jobs": len(jobs), "pool": len(pool), "slots": args.slots},
        "served": st, "served_by_kind": by_kind, "served_saturated": sat,
        "plain_per_job": dict(pst, jobs=npl, results_equal=same),
        "speedup_saturated_vs_plain": round(sat["jobs_per_s"] / pst["jobs_per_s"], 2),
        "oracle": oracle,
        "roofline": None,
        "cpu_baseline": None,
    }


def batch_sizes(nmesh, seed=2024):
    """Cells per side of the batch's cube meshes: log-uniform in [8, 96]."""
    import numpy as np
    rng = np.random.default_rng(seed)
    return np.round(np.exp(rng.uniform(np.log(8), np.log(96), nmesh))).astype(int)


def run_plain_per_mesh(args, ms, mn, mean, dt):
    """The plain per-mesh path on the same meshes: one handle per mesh,
    GmlxTetQuality + GmlxBallMin + GmlReduceVector min and sum per mesh;
    its per-mesh results must equal the batch's."""
    import numpy as np
    import gmlib_amd as G
    M = len(ms)
    t0 = time.time()
    hs = []
    for crd_m, ref_m, tet_m in ms:
        gm = G.Gml(0)
        gm.set_mesh(crd_m, ref_m, tet_m)
        q = gm.solution(G.GmlTetrahedra, 1, G.GmlFlt, "qal")
        vq = gm.solution(G.GmlVertices, 1, G.GmlFlt, "vqal")
        hs.append((gm, q, vq))
    red = C.c_double()

    def plain_step():
        out = np.empty((M, 2))
        for i, (gm, q, vq) in enumerate(hs):
            if gm.L.GmlxTetQuality(gm.h, q) != 1 or gm.L.GmlxBallMin(gm.h, q, vq) != 1:
                raise RuntimeError("plain path launch failed")
            gm.L.GmlReduceVector(gm.h, q, G.GmlMin, C.byref(red))
            out[i, 0] = red.value
            gm.L.GmlReduceVector(gm.h, q, G.GmlSum, C.byref(red))
            out[i, 1] = red.value / len(ms[i][2])
        return out

    plain_step()
    log("plain path setup (%d handles, upload, ball build): %.1fs" % (M, time.time() - t0))
    for _ in range(max(1, args.warmup - 1)):
        ref_out = plain_step()
    psteps = max(3, args.steps // 4)
    t0 = time.perf_counter()
    for _ in range(psteps):
        ref_out = plain_step()
    pdt = (time.perf_counter() - t0) / psteps
    min_equal = bool(np.array_equal(ref_out[:, 0].astype(np.float32), mn))
    mean_err = float(np.max(np.abs(ref_out[:, 1] - mean) / np.abs(mean)))
    log("plain: %.3f ms per step, %.0f meshes/s; batched / plain = %.2fx; per-mesh min equal %s, mean max rel diff %.2e"
        % (pdt * 1e3, M / pdt, pdt / dt, min_equal, mean_err))
    for gm, _, _ in hs:
        gm.close()
    if not min_equal or mean_err > 1e-12:
        raise RuntimeError("batched and plain per-mesh results differ")
    return {"ms_per_step": round(pdt * 1e3, 3), "meshes_per_s": round(M / pdt, 1),
            "speedup_batched": round(pdt / dt, 2), "min_equal": min_equal, "mean_max_rel_diff": mean_err}


def run_batch(args):
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    import gmlib_amd as G

    sizes = batch_sizes(args.nmesh)
    t0 = time.time()
    with ThreadPoolExecutor(16) as ex:
        ms = list(ex.map(lambda i: G.cube_mesh(int(sizes[i]), 0.3, 7000 + i), range(len(sizes))))
    nver, ntet, crd, tet = G.concat_meshes(ms)
    V, T, M = int(nver.sum()), int(ntet.sum()), len(ms)
    log("batch: %d meshes, n in [%d, %d], V=%d T=%d, generated in %.1fs" % (M, sizes.min(), sizes.max(), V, T, time.time() - t0))
    alg = 36 * T + 20 * V

    g = new_handle(args)
    L, h = g.L, g.h
    t0 = time.time()
    b = L.GmlxNewBatch(h, M, G.ptr(nver), G.ptr(ntet), G.ptr(crd), G.ptr(tet))
    if b <= 0:
        raise RuntimeError("GmlxNewBatch failed")
    t_new = time.time() - t0
    t0 = time.time()
    tv, pad = C.c_int(), C.c_int()
    if args.shape:
        tv.value, pad.value = (int(x) for x in args.shape.split("/"))
        if L.GmlxBatchSetShape(h, b, tv.value, pad.value) != 1:
            raise RuntimeError("GmlxBatchSetShape failed")
    elif L.GmlxBatchTune(h, b, C.byref(tv), C.byref(pad)) != 1:
        raise RuntimeError("GmlxBatchTune failed")
    t_tune = time.time() - t0
    tab = (C.c_double * 27)()
    nt = L.GmlxGetBatchTuning(h, b, tab)
    tuning = {"%d/%d" % (tab[3 * i], tab[3 * i + 1]): round(tab[3 * i + 2], 1) for i in range(nt)}
    info = (C.c_longlong * 12)()
    L.GmlxGetBatchInfo(h, b, info)
    log("batch upload %.1fs, tune + plan + graph %.1fs: TV %d, Pad %d; tuner (TV/Pad: us on the sample) %s; "
        "%d tiles, tile tets / T %.3f, tile vertices / V %.3f, plan %.2f GB, %d work-groups"
        % (t_new, t_tune, tv.value, pad.value, tuning, info[3], info[7] / T, info[8] / V, info[6] / 1e9, info[9]))

    if args.copy_calib:
        vc = L.GmlNewVector(h, 1 << 22, 4, None, G.GmlFlt)
        one = C.c_double(1.0)
        for _ in range(args.copy_calib):
            L.GmlScaleVec(h, vc, C.byref(one))
        L.GmlxSync(h)
        L.GmlxFreeVector(h, vc)
        log("copy_calib: %d GmlScaleVec launches (k_scale) of %d bytes read + %d bytes written"
            % (args.copy_calib, 1 << 26, 1 << 26))

    mn = np.empty(M, np.float32)
    mean = np.empty(M, np.float64)
    for _ in range(args.warmup):
        L.GmlxBatchQuality(h, b, G.ptr(mn), G.ptr(mean))
    cnt = C.c_longlong()
    k0 = L.GmlxGetBuiltinRunTime(h, G.GmlxBuiltinBatch, C.byref(cnt))
    c0 = cnt.value
    t0 = time.perf_counter()
    for _ in range(args.steps):
        if L.GmlxBatchQuality(h, b, G.ptr(mn), G.ptr(mean)) != 1:
            raise RuntimeError("GmlxBatchQuality failed")
    dt = (time.perf_counter() - t0) / args.steps
    kavg = (L.GmlxGetBuiltinRunTime(h, G.GmlxBuiltinBatch, C.byref(cnt)) - k0) / (cnt.value - c0)
    kgbs = alg / kavg / 1e9
    log("batched: %.3f ms per step (graph %.3f ms), %.0f meshes/s, %.2f Gtets/s, %.1f GB/s algorithmic (%.1f%% of peak)"
        % (dt * 1e3, kavg * 1e3, M / dt, T / dt / 1e9, kgbs, kgbs / HBM_PEAK_GBS * 100))

    plain = None
    if args.plain:
        plain = run_plain_per_mesh(args, ms, mn, mean, dt)

    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_batch512.json")
    if args.nmesh == 512 and os.path.exists(tf):
        traffic = json.load(open(tf)).get("batch_quality")
    # the kernel's own compulsory bytes: its tile plan (vertex lists, tile
    # tets as 4 x uint16, ball slabs as uint16, offsets), the coordinates of
    # every tile's vertices (halo included), quality and ball minimum out,
    # the per-tile partials -- fewer than the model's 36 T + 20 V
    bi = (C.c_longlong * 12)()
    L.GmlxGetBatchInfo(h, b, bi)
    own = bi[6] + 16 * bi[8] + 4 * T + 4 * V + 16 * bi[3]
    out = {
        "metric": "HBM GB/s (% peak) per kernel; served mesh-jobs/s + p99 vs plain path",
        "value": round(alg / dt / 1e9, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded jittered Kuhn cube meshes, cells per side log-uniform in [8, 96])",
        "config": {"workload": "batched tet quality + ball min + per-mesh min/mean, %d cube meshes (%d vertices, "
                               "%d tets), one persistent kernel from a HIP graph" % (M, V, T),
                   "meshes": M, "vertices": V, "tets": T, "bytes_per_step": alg, "tile_vertices": tv.value,
                   "ball_padding": pad.value, "tuner_us": tuning, "setup_s": {"upload": round(t_new, 2),
                                                                              "tune_plan_graph": round(t_tune, 2)}},
        "meshes_per_s": round(M / dt, 1),
        "plain_per_mesh": plain,
        "roofline": {"bound": "hbm", "achieved": round(kgbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(kgbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_batch_quality<%d> (graph: + k_batch_finish)" % pad.value,
                     "bytes_per_launch": alg, "avg_us": round(kavg * 1e6, 2),
                     "own_bytes_per_launch": own, "own_frac": round(own / kavg / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic_frac": round(traffic / kavg / 1e9 / HBM_PEAK_GBS, 4) if traffic else None},
        "cpu_baseline": None,
    }
    if args.cpu_seconds > 0:
        from tests import oracle_lib as O
        # the generated quality and ball-minimum kernels built for the host
        # cores (oracle/host_kernels.py), the reductions by the oracle's C
        # loops.  Vertex balls are set-up work on both sides (the GPU builds
        # them at batch creation): built before the timed region
        from oracle import host_kernels as HK
        src = HK.quality_sources(G, os.path.join(ROOT, "gmlib_amd", "kernels"))
        hk = HK.HostKernels(HK.build(src, tag="quality"), src)
        tb = time.perf_counter()
        hqs = []
        for i in range(M):
            deg, ball, md = O.vertex_ball(ms[i][2], len(ms[i][0]), 32)
            if md > 32:
                break
            hqs.append(HK.Quality(hk, ms[i][0], ms[i][2], deg, ball))
            if time.perf_counter() - tb > args.cpu_seconds:
                break
        t0 = time.perf_counter()
        done_t = done_v = 0
        i = 0
        while i < len(hqs) and (i < 8 or time.perf_counter() - t0 < args.cpu_seconds):
            hq = hqs[i]
            hq.step()
            O.reduce(G.GmlMin, hq.qal)
            O.reduce(G.GmlSum, hq.qal)
            done_t += hq.T
            done_v += hq.V
            i += 1
        ct = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round((36 * done_t + 20 * done_v) / ct / 1e9, 2), "unit": "GB/s",
                               "cores": hk.threads, "kind": "port",
                               "sample": "the first %d meshes of the batch (%d tets): the generated tet_meanratio and "
                                         "ball_min kernels compiled for x86-64 by clang and run with OpenMP (balls "
                                         "built beforehand), min and sum by the oracle's C loops, %.2f s"
                                         % (i, done_t, ct)}
        log("cpu baseline: %s" % out["cpu_baseline"])
    g.close()
    return out


def extract_bytes(T, E, F):
    """Algorithmic bytes of one edge + face extraction: the tet table read by
    each call, the edge table (2 nodes + ref) and triangle table (int4 row
    + ref) written."""
    return 32 * T + 12 * E + 20 * F


def run_extract(args):
    import numpy as np
    import gmlib_amd as G

    n = args.n
    crd, ref, tet = G.cube_mesh(n, 0.2, 2024, True)
    V, T = len(crd), len(tet)
    g = new_handle(args)
    L, h = g.L, g.h
    g.set_mesh(crd, ref, tet)
    L.GmlxSync(h)

    def drop(typ):
        idx = C.c_int(0)
        if L.GmlGetMeshInfo(h, typ, None, C.byref(idx)):
            L.GmlFreeData(h, idx.value)

    def step():
        drop(G.GmlEdges)
        drop(G.GmlTriangles)
        L.GmlxSync(h)
        t0 = time.perf_counter()
        ne = L.GmlExtractEdges(h)
        t1 = time.perf_counter()
        nf = L.GmlExtractFaces(h)
        t2 = time.perf_counter()
        if ne <= 0 or nf <= 0:
            raise RuntimeError("extraction failed (%d edges, %d faces)" % (ne, nf))
        return ne, nf, t1 - t0, t2 - t1

    for _ in range(args.warmup):
        step()
    te = tf = 0.0
    for _ in range(args.steps):
        ne, nf, a, b = step()
        te += a
        tf += b
    E = 3 * n * (n + 1) ** 2 + 3 * n * n * (n + 1) + n ** 3
    F = (4 * T + 12 * n * n) // 2
    if ne != E or nf != F:
        raise RuntimeError("extraction counts %d, %d differ from the cube's %d edges, %d faces" % (ne, nf, E, F))
    dt = (te + tf) / args.steps
    alg = extract_bytes(T, E, F)
    gbs = alg / dt / 1e9
    log("extract: %d tets -> %d edges in %.2f ms, %d triangles in %.2f ms; %.1f GB/s algorithmic, %.2f Gtets/s"
        % (T, E, te / args.steps * 1e3, F, tf / args.steps * 1e3, gbs, T / dt / 1e9))
    out = {
        "metric": "HBM GB/s (% peak) per kernel; served mesh-jobs/s + p99 vs plain path",
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded jittered Kuhn cube mesh)",
        "config": {"workload": "GmlExtractEdges + GmlExtractFaces, %d^3-cell Kuhn cube (%d vertices, %d tets -> "
                               "%d edges, %d triangles)" % (n, V, T, E, F),
                   "vertices": V, "tets": T, "edges": E, "triangles": F, "bytes_per_step": alg,
                   "edges_ms": round(te / args.steps * 1e3, 3), "faces_ms": round(tf / args.steps * 1e3, 3)},
        "tets_per_s": round(T / dt, 1),
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "whole extraction (vertex-ball first appearances, flag compaction, row writes)",
                     "bytes_per_launch": alg, "avg_us": round(dt * 1e6, 1)},
        "cpu_baseline": None,
    }
    # bytes moved per step by all the step's kernels, from the committed
    # rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this workload
    trf = os.path.join(ROOT, "profiles", "traffic_extract256.json")
    if n == 256 and os.path.exists(trf):
        out["roofline"]["traffic"] = json.load(open(trf)).get("extract_step")
    g.close()
    if args.cpu_seconds > 0:
        from tests import oracle_lib as O
        m = 8
        best = None
        while True:
            _, _, ct = G.cube_mesh(m, 0.2, 2024, True)
            t0 = time.perf_counter()
            e, _ = O.extract_items(2, [(G.GmlTetrahedra, ct)], tabsiz=len(ct))
            f, _ = O.extract_items(3, [(G.GmlTetrahedra, ct)])
            c = time.perf_counter() - t0
            best = (m, len(ct), len(e), len(f), c)
            if c > args.cpu_seconds / 4 or m >= n:
                break
            m = int(m * 1.6)
        m, ct_n, ce, cf, c = best
        out["cpu_baseline"] = {"value": round(extract_bytes(ct_n, ce, cf) / c / 1e9, 3), "unit": "GB/s",
                               "cores": 1, "kind": "port",
                               "sample": "%d^3-cell cube (%d tets -> %d edges, %d triangles): the oracle's restatement "
                                         "of the reference hash (GetItmNod / CalHshKey / AddHsh / GetHsh), one "
                                         "thread, %.2f s" % (m, ct_n, ce, cf, c)}
        log("cpu baseline: %s" % out["cpu_baseline"])
    return out


def heat_bytes(T, F, V):
    """Algorithmic bytes of one heat iteration, each datum once: face values
    (shell slots + degree + ref 16F, cell values 4T, out 4F), cell gradients
    (connectivity 16T, coordinates 16V, downlink 16T, face values 4F, out
    16T), face gradients (shell + ref 16F, cell gradients 16T, out 16F), flux
    balance (16T + 16V + 16T, face gradients 16F, out 4T), update (12T), L2
    residual (4T)."""
    return 120 * T + 72 * F + 32 * V


def run_heat(args):
    import numpy as np
    import gmlib_amd as G
    from gmlib_amd import heat as H

    n = args.n if args.n != 256 else 128
    t0 = time.time()
    crd, vref, tet, tri, side = H.heat_cube(n, 0.2, 11)
    s = H.HeatSolver(crd, vref, tet, tri, side)
    T, F, V = len(tet), s.nface, len(crd)
    log("heat: %d^3 cube, %d tets, %d faces, set-up (mesh, faces, links, compile, dt) %.1fs"
        % (n, T, F, time.time() - t0))
    s.iterate(max(args.warmup, 1))
    for _ in range(100):                 # until every kernel's launch-shape tuner has settled
        if all(tuned(s.g.L, s.g.h, k) for k in (s.k_face, s.k_grad, s.k_fgrad, s.k_bal, s.k_step)):
            break
        s.iterate(1)
    kern = {"heat_face_value": s.k_face, "heat_cell_gradient": s.k_grad, "heat_face_gradient": s.k_fgrad,
            "heat_balance": s.k_bal, "heat_step": s.k_step}
    k0 = {k: s.g.kernel_time(v) for k, v in kern.items()}
    s.g.L.GmlxSync(s.g.h)
    t0 = time.perf_counter()
    res = s.iterate(args.steps)
    s.g.L.GmlxSync(s.g.h)
    dt = (time.perf_counter() - t0) / args.steps
    kt = {k: (s.g.kernel_time(v) - k0[k]) / args.steps for k, v in kern.items()}
    kb = {"heat_face_value": 20 * F + 4 * T, "heat_cell_gradient": 48 * T + 16 * V + 4 * F,
          "heat_face_gradient": 32 * F + 16 * T, "heat_balance": 36 * T + 16 * V + 16 * F, "heat_step": 12 * T}
    dom = max(kt, key=kt.get)
    alg = heat_bytes(T, F, V)
    log("heat: %.3f ms per iteration (%.1f GB/s of the byte model); kernels (us): %s; residual %.3e -> %.3e"
        % (dt * 1e3, alg / dt / 1e9, {k: round(v * 1e6, 1) for k, v in kt.items()}, res[0], res[-1]))
    kgbs = kb[dom] / kt[dom] / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_heat128.json")
    if n == 128 and os.path.exists(tf):
        traffic = json.load(open(tf)).get(dom)
    out = {
        "metric": "HBM GB/s (% peak) per kernel; served mesh-jobs/s + p99 vs plain path",
        "value": round(alg / dt / 1e9, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded jittered Kuhn cube mesh, boundary triangles refs 1-6 by side)",
        "config": {"workload": "cell-centered heat equation (HeatCellCentered), %d^3-cell Kuhn cube (%d tets, "
                               "%d faces, %d vertices), one explicit iteration + L2 residual per step" % (n, T, F, V),
                   "tets": T, "faces": F, "vertices": V, "bytes_per_step": alg,
                   "kernel_us": {k: round(v * 1e6, 2) for k, v in kt.items()},
                   "kernel_frac": {k: round(kb[k] / v / 1e9 / HBM_PEAK_GBS, 4) for k, v in kt.items()},
                   "kernel_form": {k: FORMS.get(s.g.L.GmlxGetKernelForm(s.g.h, v), "?") for k, v in kern.items()}},
        "iterations_per_s": round(1 / dt, 1),
        "roofline": {"bound": "hbm", "achieved": round(kgbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(kgbs / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom,
                     "bytes_per_launch": kb[dom], "avg_us": round(kt[dom] * 1e6, 2)},
        "cpu_baseline": None,
    }
    s.close()
    if args.cpu_seconds > 0:
        from tests import oracle_lib as O
        m = min(n, 48)
        c, vr, tt, tr, sd = H.heat_cube(m, 0.2, 11)
        full, fref = O.extract_items(3, [(G.GmlTetrahedra, tt)], tr, sd)
        O.heat_iterations(c, tt, full, fref, H.DEFAULT_BC, 1)
        k = 1
        while True:
            t0 = time.perf_counter()
            O.heat_iterations(c, tt, full, fref, H.DEFAULT_BC, k)
            ct = time.perf_counter() - t0
            if ct > args.cpu_seconds / 3 or k >= 64:
                break
            k *= 2
        per = ct / k
        out["cpu_baseline"] = {"value": round(heat_bytes(len(tt), len(full), len(c)) / per / 1e9, 3), "unit": "GB/s",
                               "cores": 1, "kind": "port",
                               "sample": "%d iterations of the numpy FP64 restatement (links and geometry set up "
                                         "once per call) on a %d^3 cube (%d tets), %.2f s" % (k, m, len(tt), ct)}
        log("cpu baseline: %s" % out["cpu_baseline"])
    return out


def bandwidth_models(V, T):
    """per kernel: (algorithmic bytes, each datum once) and the reference
    example's own figures (gpu_bandwidth.c:151-176: direct 32V + 32T for the
    vertex and tet kernels together; indirect 96T raw, 16V + 32T unique;
    ragged 460V raw, 32V + 16T unique)"""
    alg = {"direct_vertices": 32 * V, "direct_tets": 32 * T, "indirect_tets": 16 * V + 32 * T,
           "vertices_ball": 32 * T + 20 * V}       # ball slots 16T, degree 4V, TetDat 16T, VerDat 16V
    return alg


def bandwidth_pass(args, G, crd, ref, tet, steps, warmup, tag=""):
    """the four GpuBandwidth kernels on one handle: per-kernel average
    seconds (HIP events on the library's stream), GmlEvaluateNumbering;
    tag is appended to the kernel names (profiles tell the passes apart)"""
    R, W = G.GmlReadMode, G.GmlWriteMode
    g = new_handle(args)
    try:
        vi, ti = g.set_mesh(crd, ref, tet)
        score = g.L.GmlEvaluateNumbering(g.h)
        vdat = g.solution(G.GmlVertices, 1, G.GmlFlt4, "VerDat")
        tdat = g.solution(G.GmlTetrahedra, 1, G.GmlFlt4, "TetDat")
        ks = {"direct_vertices": g.compile(body("bw_direct_vertices"), "direct_vertices" + tag, G.GmlVertices,
                                           [(vi, R, 0), (vdat, W, 0)]),
              "direct_tets": g.compile(body("bw_direct_tets"), "direct_tets" + tag, G.GmlTetrahedra,
                                       [(ti, R, 0), (tdat, W, 0)]),
              "indirect_tets": g.compile(body("bw_indirect_tets"), "indirect_tets" + tag, G.GmlTetrahedra,
                                         [(vi, R, 0), (tdat, W, 0)]),
              "vertices_ball": g.compile(body("bw_vertices_ball"), "vertices_ball" + tag, G.GmlVertices,
                                         [(vdat, W, 0), (tdat, R, 0)])}
        t, forms = {}, {}
        for name, k in ks.items():           # the example's order: direct, indirect, ragged
            # past the launch tuner's trials of every shape, then the warm-up
            for i in range(200):
                if i >= warmup and tuned(g.L, g.h, k):
                    break
                if g.L.GmlLaunchKernel(g.h, k) != 1:
                    raise RuntimeError("launch failed")
            g.L.GmlxSync(g.h)
            t0 = g.kernel_time(k)
            for _ in range(steps):
                g.L.GmlLaunchKernel(g.h, k)
            g.L.GmlxSync(g.h)
            t[name] = (g.kernel_time(k) - t0) / steps
            forms[name] = FORMS.get(g.L.GmlxGetKernelForm(g.h, k), "?")
        return t, score, forms
    finally:
        g.close()


def run_bandwidth(args):
    import numpy as np
    import gmlib_amd as G

    n = args.n
    t0 = time.time()
    crd0, ref0, tet0 = G.cube_mesh(n, 0.2, 2024, False)
    V, T = len(crd0), len(tet0)
    rng = np.random.default_rng(7)
    pv = rng.permutation(V)
    inv = np.empty_like(pv)
    inv[pv] = np.arange(V)
    crdr = np.ascontiguousarray(crd0[pv])
    refr = np.ascontiguousarray(ref0[pv])
    tetr = np.ascontiguousarray(inv[tet0[rng.permutation(T)]].astype(np.int32))
    del crd0, ref0, tet0, pv, inv
    log("bandwidth: %d^3 cube (V=%d, T=%d), random numbering in %.1fs" % (n, V, T, time.time() - t0))
    alg = bandwidth_models(V, T)
    res = {}
    tr, sr, fr = bandwidth_pass(args, G, crdr, refr, tetr, args.steps, args.warmup, "_random")
    res["random"] = (tr, sr, fr)
    crdh, refh, teth = crdr.copy(), refr.copy(), tetr.copy()
    t0 = time.time()
    if G.lib().GmlxHilbertRenumber(V, G.ptr(crdh), G.ptr(refh), T, 4, G.ptr(teth), None, None, None, 0) != 1:
        raise RuntimeError("GmlxHilbertRenumber failed")
    log("bandwidth: GPU Hilbert renumbering %.2fs" % (time.time() - t0))
    del crdr, refr, tetr
    th, sh, fh = bandwidth_pass(args, G, crdh, refh, teth, args.steps, args.warmup)
    res["hilbert"] = (th, sh, fh)

    def summary(t, score, forms):
        k = {name: {"avg_us": round(v * 1e6, 2), "alg_GBs": round(alg[name] / v / 1e9, 1),
                    "frac": round(alg[name] / v / 1e9 / HBM_PEAK_GBS, 4), "form": forms[name]}
             for name, v in t.items()}
        direct = t["direct_vertices"] + t["direct_tets"]
        refm = {"direct_GBs": round((32 * V + 32 * T) / direct / 1e9, 1),
                "indirect_GBs": round(96 * T / t["indirect_tets"] / 1e9, 1),
                "indirect_unique_GBs": round((16 * V + 32 * T) / t["indirect_tets"] / 1e9, 1),
                "ragged_GBs": round(460 * V / t["vertices_ball"] / 1e9, 1),
                "ragged_unique_GBs": round((32 * V + 16 * T) / t["vertices_ball"] / 1e9, 1)}
        return {"kernels": k, "reference_byte_model": refm, "evaluate_numbering": round(score, 4),
                "GBs": round(sum(alg.values()) / sum(t.values()) / 1e9, 1)}

    outr, outh = summary(*res["random"]), summary(*res["hilbert"])
    log("random numbering: %s" % outr)
    log("hilbert numbering: %s" % outh)
    dom = max(th, key=th.get)
    kg = alg[dom] / th[dom] / 1e9
    step = sum(th.values())
    out = {
        "metric": "HBM GB/s (% peak) per kernel; served mesh-jobs/s + p99 vs plain path",
        "value": outh["GBs"],
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded jittered Kuhn cube mesh, randomly renumbered, then Hilbert-renumbered on the GPU)",
        "config": {"workload": "GpuBandwidth access patterns through the generated kernels (direct vertex and tet "
                               "copies, indirect tet -> vertex gather, ragged vertex -> tet ball gather), "
                               "%d^3-cell Kuhn cube (%d vertices, %d tets)" % (n, V, T),
                   "vertices": V, "tets": T, "bytes_per_step": sum(alg.values())},
        "hilbert": outh,
        "random": outr,
        "roofline": {"bound": "hbm", "achieved": round(kg, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(kg / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dom,
                     "bytes_per_launch": alg[dom], "avg_us": round(th[dom] * 1e6, 2)},
        "cpu_baseline": None,
    }
    tf = os.path.join(ROOT, "profiles", "traffic_bandwidth256.json")
    if n == 256 and os.path.exists(tf):
        out["roofline"]["traffic"] = json.load(open(tf)).get(dom)
    if args.cpu_seconds > 0:
        # the oracle's OpenMP loops of the same patterns on the Hilbert-numbered mesh
        from tests import oracle_lib as O
        OL = O.lib()
        deg, ball, _ = O.vertex_ball(teth, V, 32)
        tdat = np.empty((T, 4), np.float32)
        vdat = np.empty((V, 4), np.float32)
        loops = [("direct_vertices", lambda: np.copyto(vdat, crdh)),
                 ("direct_tets", lambda: OL.orc_bw_direct_tets(T, O.p(teth), O.p(tdat))),
                 ("indirect_tets", lambda: OL.orc_bw_indirect_tets(T, O.p(teth), O.p(crdh), O.p(tdat))),
                 ("vertices_ball", lambda: OL.orc_bw_vertices_ball(V, O.p(deg), O.p(ball), 32, O.p(tdat), O.p(vdat)))]
        ct, ns = 0.0, 0
        while ns < 2 or ct < args.cpu_seconds:
            for name, f in loops:
                t0 = time.perf_counter()
                f()
                ct += time.perf_counter() - t0
            ns += 1
        out["cpu_baseline"] = {"value": round(ns * sum(alg.values()) / ct / 1e9, 2), "unit": "GB/s",
                               "cores": OL.orc_num_threads(), "kind": "port",
                               "sample": "%d passes of the four patterns over the same Hilbert-numbered %d^3 mesh, "
                                         "oracle C + OpenMP (direct vertex copy: numpy), %.2f s" % (ns, n, ct)}
        log("cpu baseline: %s" % out["cpu_baseline"])
    return out


if __name__ == "__main__":
    main()
