/*
 * gmlx.h -- extensions of the MI355X-native GMlib engine that have no
 * counterpart in the reference interface.  Plain C ABI: integers, doubles,
 * pointers and sizes only.  Each entry point states who calls it.
 */

#ifndef GMLX_H
#define GMLX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* --- handles ---------------------------------------------------------------- */

/* A handle with no device, for inspection and compile checks on machines
   without a GPU: data live in host memory, links are built on the host,
   kernels are generated and compiled for gfx950 (GmlxGetKernelSource), but
   GmlLaunchKernel returns -1 and GmlReduceVector -6.  Release with GmlStop. */
size_t GmlxInitHostOnly(void);

/* --- reductions without allocation ---------------------------------------- */

/* GmlReduceVector (sources/gmlib3.c:3229-3332) with nothing allocated at
   call time, for callers that cannot allocate (code under HIP graph capture,
   tight serving loops, harnesses that forbid hipMalloc).  No counterpart in
   the reference.  Same arguments, result and error codes as
   GmlReduceVector, from the same kernels in the same order (so the same
   double); the work-group partials go to a file-scope device array of
   16,384 partials + the result (131,080 bytes per GPU, held until the
   process ends) instead of the handle's scratch, and no timing events are
   recorded.  The grid is capped at those 16,384 work-groups, so every
   vector length fits.  Calls that overlap on one GPU (two handles' streams
   at once) race on that array; calls in sequence do not, since each launch
   writes every partial it reads. */
int    GmlxReduceVectorNoAlloc(size_t GmlIdx, int DatIdx, int RedOpp, double *res);

/* --- FFI-friendly forms of variadic entry points ------------------------- */

/* GmlCompileKernel with the (DatIdx, flags, LnkIdx) triplets in an array of
   3*NmbTyp ints, for callers that cannot build a C varargs call (ctypes,
   cgo, JNI).  Same result and error codes as GmlCompileKernel. */
int    GmlxCompileKernelv(size_t GmlIdx, const char *KrnSrc, const char *PrcNam,
                          int MshTyp, int NmbTyp, const int *triplets);

/* LDS staging of uplink data in generated kernels (default on): read-only,
   single-item data reached through a ball (e.g. tet data read by a vertex
   kernel) are loaded once per work-group tile into LDS with coalesced loads
   and read from there; results are identical either way.  Applies to
   kernels the handle compiles afterwards; returns the previous setting. */
int    GmlxSetStaging(size_t GmlIdx, int on);

/* Generated HIP source of a compiled kernel (NUL-terminated, owned by the
   library, valid until GmlStop); NULL if the index is invalid.  The paired
   form (two rows per thread, the reads of both issued before either body;
   built for kernels without uplinks) has its own source, NULL if none. */
const char *GmlxGetKernelSource(size_t GmlIdx, int KrnIdx);
const char *GmlxGetKernelPairSource(size_t GmlIdx, int KrnIdx);
/* the element-tile form's source (see GmlxGetKernelForm), NULL if none */
const char *GmlxGetKernelTileSource(size_t GmlIdx, int KrnIdx);

/* Group size and XCD remap flag the launch-shape tuner settled on for a
   kernel (0 in both while it is still tuning). */
int    GmlxGetKernelShape(size_t GmlIdx, int KrnIdx, int *GrpSiz, int *Remap);

/* Rows per thread of a kernel's launches: 1, 2 (its paired form) or, in the
   element-tile form, the tile's rows per work-item (512 / group size); 0
   while the tuner is still choosing; -1 for an invalid index. */
int    GmlxGetKernelRows(size_t GmlIdx, int KrnIdx);

/* Form of a kernel's launches: 1 one row per work-item, 2 the paired form,
   3 the element-tile form (an element loop reading vertex data through its
   own table: per work-group a tile of 512 elements whose distinct vertices
   are loaded once into LDS; built when the body never names the table row,
   used when every tile of the table reaches at most 256 vertices); 0 while
   the tuner is still choosing; -1 for an invalid index. */
int    GmlxGetKernelForm(size_t GmlIdx, int KrnIdx);

/* Element tile plan of element kind typ (built if needed; see
   GmlxGetKernelForm), 6 values: {tile rows, tiles, sum of the tiles' vertex
   counts, largest tile, index words per element, plan bytes}.  Returns 1,
   or 0 if the kind has no plan. */
int    GmlxGetTilePlanInfo(size_t GmlIdx, int typ, long long *info);

/* Pin the launch shape of a kernel (disables its tuner): group size 64, 128
   or 256 (generated kernels are compiled for at most 256 work-items), XCD
   remap, rows per thread (2 needs the paired form; -1 selects the
   element-tile form).  Returns 1, or 0 if the shape is not available for
   the kernel.  A kernel whose body calls a work-item function
   (get_global_id ...) runs in plain block order, one row per work-item, so
   that get_global_id(0) is its row as in OpenCL: no remap, no other form. */
int    GmlxSetKernelShape(size_t GmlIdx, int KrnIdx, int GrpSiz, int Remap, int Rows);

/* --- device residency ---------------------------------------------------- */

/* Device address and size in bytes of a data item (for callers that keep
   their own device buffers next to the library's, e.g. torch tensors). */
void  *GmlxGetDevicePointer(size_t GmlIdx, int DatIdx, size_t *bytes);

/* Wait for every command issued on the handle's stream. */
int    GmlxSync(size_t GmlIdx);

/* The handle's HIP stream (hipStream_t), for timing or interop. */
void  *GmlxGetStream(size_t GmlIdx);

/* Mark a data item's device contents as changed behind the library's back
   (the next GmlGetDataLine/Block downloads it again). */
int    GmlxInvalidateHost(size_t GmlIdx, int DatIdx);

/* --- topology ------------------------------------------------------------ */

/* Build (if needed) the SrcTyp -> DstTyp link that GmlCompileKernel would
   build, and return 1.  On a device handle every kind is built on the GPU:
   vertex balls, edge and face shells, downlinks (element -> its edges or
   faces). */
int    GmlxBuildLink(size_t GmlIdx, int SrcTyp, int DstTyp);

/* Data index of the SrcTyp -> DstTyp link table (a downlink, or the
   neighbours of GmlSetNeighbours when SrcTyp == DstTyp), 0 if none; its
   lines are read with GmlGetDataLine. */
int    GmlxGetLinkData(size_t GmlIdx, int SrcTyp, int DstTyp);

/* Copy an uplink out in the reference's row-major form: deg[NmbSrc] and
   ball[NmbSrc * width] with entries (element << 4) | local index in
   descending element order, 0-padded.  width must be >= the max degree. */
int    GmlxGetUplinkTable(size_t GmlIdx, int SrcTyp, int DstTyp, int *deg, int *ball, int width);

/* --- block vectors and solver ------------------------------------------- */

/* Copy a vector made by GmlNewVector in or out (NmbLin * BlkSiz packed
   values of its float type).  No counterpart: the reference offers no way to
   read a vector back. */
int    GmlxSetVector(size_t GmlIdx, int VecIdx, const void *val);
int    GmlxGetVector(size_t GmlIdx, int VecIdx, void *val);
int    GmlxFreeVector(size_t GmlIdx, int VecIdx);
int    GmlxFreeMatrix(size_t GmlIdx, int MatIdx);

/* Layout of a block matrix made by GmlNewMatrix, 6 values: {block rows,
   blocks, block size, SELL-64 slabs, slabs with 16-bit columns, blocks read
   through a 16-bit column}.  Each row's blocks whose column lies within
   +-32767 of the row come first (the caller's order otherwise kept); a
   64-row slab's first kn slots, kn its rows' smallest count of such blocks,
   also hold their columns as 16-bit offsets, which the products read instead
   of the 32-bit columns.  Returns 1, or 0 for an invalid index. */
int    GmlxGetMatrixInfo(size_t GmlIdx, int MatIdx, long long *info);

/* Dot product of two FP32 vectors with FP64 accumulation. */
int    GmlxDotVec(size_t GmlIdx, int VecA, int VecB, double *res);

/* Block-Jacobi preconditioned conjugate gradient on A x = rhs for an SPD
   FP32 block matrix (block size 1, 3 or 4).  refine = 0: one FP32 solve to
   the relative tolerance tol; refine = 1: iterative refinement, x held in
   FP64, residual rhs - A x accumulated in FP64, FP32 inner solves, until the
   FP64 relative residual is <= tol.  SolIdx may be an FP32 or FP64 vector.
   *relres = |rhs - A x| / |rhs| evaluated in FP64, *iters = inner iterations.
   Returns 1 or a negative error. */
int    GmlxSolveCG(size_t GmlIdx, int MatIdx, int RhsIdx, int SolIdx, double tol, int maxit,
                   int refine, double *relres, int *iters);

/* --- hand-written mesh kernels --------------------------------------------- */

/* HIP kernels written for gfx950 that compute exactly what the generated
   kernels of the bodies in gmlib_amd/kernels compute (tet_meanratio.cl,
   ball_min.cl, smooth_scatter.cl + smooth_gather.cl), on the handle's vertex
   and tet tables and vertex->tet ball (built on first use).  For callers that
   run those workloads at scale; launches are asynchronous on the handle's
   stream. */
enum gmlx_builtin { GmlxBuiltinQuality, GmlxBuiltinBallMin, GmlxBuiltinScatter, GmlxBuiltinGather,
                    GmlxBuiltinSweep, GmlxBuiltinSpMV, GmlxBuiltinCgSpMV, GmlxBuiltinCgUpdate,
                    GmlxBuiltinBatch };

/* mean-ratio quality of every tet into QalIdx (a GmlFlt tet solution), from
   the x, y, z of the vertices (the fourth coordinate slot, 0 whenever the
   coordinates come through GmlSetDataLine/Block, is not read) */
int    GmlxTetQuality(size_t GmlIdx, int QalIdx);
/* per-vertex minimum of a GmlFlt tet field over the vertex ball */
int    GmlxBallMin(size_t GmlIdx, int TetValIdx, int VerOutIdx);
/* nsweep Jacobi sweeps of barycentre smoothing of the vertex coordinates
   (vertices with ref != 0 stay fixed).  Modes (GmlxSetSmoothMode):
   2 (default) operator form: one kernel per sweep over 256-vertex tiles,
     1-ring coordinates staged in LDS, x_v += omega (sum_u c_vu x_u /
     (4 deg v) - x_v) with c_vu the number of ball tets sharing u; equal to
     the two-kernel result within FP32 round-off;
   1 fused exact: one kernel per sweep, barycentres formed in LDS, results
     bit-identical to the scatter + gather kernels;
   0 two kernels: barycentre scatter into BaryIdx (a GmlFlt4 tet solution),
     then ball gather.
   Modes 1 and 2 build a tile plan on first use and double-buffer the
   coordinates: after a sweep the vertex table's device address is the other
   buffer (query it again with GmlxGetDevicePointer). */
int    GmlxSmoothSweeps(size_t GmlIdx, int BaryIdx, int nsweep, double omega);
int    GmlxSetSmoothMode(size_t GmlIdx, int mode);
/* tile plan of the current sweep mode (info: 9 values): {tile size, tiles,
   sum of tile vertices, sum of tile tets, max vertices per tile, max tets per
   tile, plan bytes, total 1-ring slots (operator form), mode} */
int    GmlxGetSweepPlanInfo(size_t GmlIdx, long long *info);
/* launch shape of a built-in kernel on one handle: GrpSiz a
   multiple of 64, up to 1024 (quality, ball-min, scatter, gather, sweep: the
   tile of the operator-form sweep; the exact fused sweep takes up to 256) or
   256 (SpMV, CG kernels); Remap 1 gives each XCD a contiguous block range */
int    GmlxSetBuiltinShape(size_t GmlIdx, int which, int GrpSiz, int Remap);
/* items per thread of a built-in kernel that takes several: the quality
   kernel (GmlxBuiltinQuality) takes 1, 2 or 4 tets per thread, or 0 (the
   default, work-groups of 128): tiles of 512 tets whose distinct vertices
   are loaded once into LDS (the tet table's tile plan, built by the second
   call on an unchanged table; until then, and without a plan, 2 tets per
   thread) */
int    GmlxSetBuiltinItems(size_t GmlIdx, int which, int PerThread);
/* summed device time (s) and launch count of a built-in kernel on a handle
   (GmlxBuiltinSpMV: the block SpMV of GmlMultMatVec; GmlxBuiltinCgSpMV and
   GmlxBuiltinCgUpdate: the two kernels of a GmlxSolveCG iteration) */
double GmlxGetBuiltinRunTime(size_t GmlIdx, int which, long long *launches);

/* --- synthetic meshes ------------------------------------------------------ */

/* Cube [0,1]^3 split in n^3 cells of 6 positively oriented (Kuhn) tets.
   crd: (n+1)^3 * 4 floats (x,y,z,0); ref: (n+1)^3 ints (1 on the boundary,
   0 inside); tet: 6 n^3 * 4 ints.  Interior vertices are jittered by a
   seeded uniform offset in [-jitter*h, jitter*h] per axis, boundary vertices
   only along their face.  order = 1 renumbers vertices and cells along a 3D
   Hilbert curve, 0 keeps lexicographic order.  Returns 1. */
int    GmlxCubeMesh(int n, double jitter, unsigned long long seed, int order,
                    float *crd, int *ref, int *tet);

/* Hilbert renumbering on the GPU (host arrays in, permuted in place):
   vertices are sorted along a 3D Hilbert curve of their bounding box (21
   bits per axis), elements (nnod vertex indices each, 0-based) along the
   same curve by their barycentre, and the element tables renumbered.  crd
   holds 4 floats per vertex; VerRef, EleRef, VerOld, EleOld may be NULL;
   VerOld / EleOld receive, for each new index, the old one.  BasDeg > 0
   also sorts the vertices by degree class: those in more than BasDeg
   elements go after all the others, each class in Hilbert order (the
   preprocessing the reference expects so that its base / high ball rows
   are two ranges, Documentation/GMlib3_en.tex:401); 0 keeps the Hilbert
   order only.  The sort is stable (deterministic).  Returns 1, or 0 on
   invalid input or GPU failure. */
int    GmlxHilbertRenumber(int NmbVer, float *crd, int *VerRef, int NmbEle, int nnod, int *ele, int *EleRef,
                           int *VerOld, int *EleOld, int BasDeg);

/* Edge-graph 3x3 vector Laplacian of a tet mesh (host, no handle): block
   CSR with one block row per vertex, columns ascending, off-diagonal block
   (i, j) = -w (I + 0.25 s s^T) with w in [0.5, 2] and s a unit vector drawn
   from a hash of (seed, min(i,j), max(i,j)), diagonal = sum of the
   incident edge blocks + shift I (SPD).  rowptr (NmbVer + 1 ints) is always
   filled; col (cap ints) and val (9 cap floats) when cap >= the block
   count.  Returns the block count, or < 0 on invalid input. */
long long GmlxEdgeLaplacian3(int NmbVer, int NmbTet, const int *tet, unsigned long long seed, double shift,
                             int *rowptr, int *col, float *val, long long cap);

/* ---- serving: a queue of mesh jobs run against device-resident meshes.
   A job names a seeded cube mesh (GmlxCubeMesh(n, 0.2, seed, Hilbert)) or a
   mesh registered with GmlxServeAddMesh, and a kind of work; up to NmbSlot jobs run at once, each on its mesh's own
   stream; meshes stay resident under an HBM budget of BudgetGB (least
   recently used idle meshes are evicted when a new one needs room); a job
   is admitted as soon as a slot is free and its mesh is resident and idle or
   fits.  Resident = 0: the plain path, every job loads its mesh into a fresh
   handle, runs and frees it, one job at a time. ---- */
enum gmlx_job { GmlxJobQuality, GmlxJobSmooth, GmlxJobCG };
size_t    GmlxServeStart(int dev, double BudgetGB, int NmbSlot, int Resident);
/* param: smooth = sweep count (20 if <= 0), cg = relative tolerance (1e-8
   if <= 0), unused for quality.  Returns a job id > 0, or -1. */
long long GmlxServeSubmit(size_t SrvIdx, int type, int n, unsigned long long seed, double param);
/* register a caller's tet mesh with the server (copied: crd 4 floats per
   vertex, ref per vertex or NULL for 0, tet 4 0-based vertex indices per
   tet); returns a mesh key > 0, or -1.  The copy stays with the server so
   the mesh can be loaded again after an eviction. */
long long GmlxServeAddMesh(size_t SrvIdx, int NmbVer, int NmbTet, const float *crd, const int *ref, const int *tet);
/* a job on a registered mesh (same types and param as GmlxServeSubmit) */
long long GmlxServeSubmitMesh(size_t SrvIdx, int type, long long MeshKey, double param);
/* forget a registered mesh (its host copy, and its device copy if resident);
   returns 1, or 0 if the key is unknown or a job on the mesh is pending or
   running */
int       GmlxServeRemoveMesh(size_t SrvIdx, long long MeshKey);
/* block until the job is done.  res[4]: quality and smooth = (min, mean
   quality), cg = (FP64 relative residual, iterations); tim[4] = (submitted,
   admitted, done) in seconds since GmlxServeStart, and 1 if the job loaded
   its mesh.  Several callers may wait on one job; each gets its result, and
   the id is retired once they return (a later wait on it returns -1).
   Returns 1, or a negative error (-7: the mesh cannot fit the budget; -8:
   the server stopped before the job was admitted). */
int       GmlxServeWait(size_t SrvIdx, long long JobIdx, double *res, double *tim);
/* 8 values: resident meshes, bytes in use, mesh loads, evictions, jobs done,
   most jobs running at once, budget bytes, peak bytes */
int       GmlxServeInfo(size_t SrvIdx, double *info);
/* stop: running jobs finish, pending jobs end with -8, callers blocked in
   GmlxServeWait return; then every resident mesh is freed */
void      GmlxServeStop(size_t SrvIdx);

/* ---- batches of meshes: many independent tet meshes kept on the device and
   processed by one persistent kernel replayed from a HIP graph.  A run
   computes the mean-ratio quality of every tet (as GmlxTetQuality), its
   minimum over every vertex ball (as GmlxBallMin) and, per mesh, the
   minimum and mean quality. ---- */
#define GmlxMaxBatch 16
/* crd: 4 floats per vertex (x, y, z, unused), meshes concatenated; tet: 4
   ints per tet, vertex indices local to their mesh, meshes concatenated.
   Returns a batch index >= 1, or 0. */
int    GmlxNewBatch(size_t GmlIdx, int NmbMsh, const int *NmbVer, const int *NmbTet,
                    const float *crd, const int *tet);
/* autotuner: time every (tile vertices TV in {256, 512, 1024}, ball padding
   Pad in {1, 2, 4}) candidate on a sample of about 1/16 of the meshes, keep
   the fastest and build its plan and graph (done by the first
   GmlxBatchQuality when not called before); *TV, *Pad receive the choice */
int    GmlxBatchTune(size_t GmlIdx, int BatIdx, int *TV, int *Pad);
/* use a given shape instead (TV a multiple of 64 in 64..1024, Pad 1, 2 or 4) */
int    GmlxBatchSetShape(size_t GmlIdx, int BatIdx, int TV, int Pad);
/* one run; MshMin / MshMean (NmbMsh entries each, either may be NULL)
   receive each mesh's minimum and mean tet quality */
int    GmlxBatchQuality(size_t GmlIdx, int BatIdx, float *MshMin, double *MshMean);
/* per-tet quality and per-vertex ball minimum of the last run, in batch
   order (either may be NULL) */
int    GmlxGetBatchField(size_t GmlIdx, int BatIdx, float *TetQal, float *VerMin);
/* 12 values: meshes, V, T, tiles, TV, Pad, plan bytes, tile tets, tile
   vertices, work-groups, graph ready, ball slabs */
int    GmlxGetBatchInfo(size_t GmlIdx, int BatIdx, long long *info);
/* tuner table: (TV, Pad, microseconds per sample run) per candidate; returns the count */
int    GmlxGetBatchTuning(size_t GmlIdx, int BatIdx, double *tab);
int    GmlxFreeBatch(size_t GmlIdx, int BatIdx);

#ifdef __cplusplus
}
#endif

#endif
