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

/* --- FFI-friendly forms of variadic entry points ------------------------- */

/* GmlCompileKernel with the (DatIdx, flags, LnkIdx) triplets in an array of
   3*NmbTyp ints, for callers that cannot build a C varargs call (ctypes,
   cgo, JNI).  Same result and error codes as GmlCompileKernel. */
int    GmlxCompileKernelv(size_t GmlIdx, const char *KrnSrc, const char *PrcNam,
                          int MshTyp, int NmbTyp, const int *triplets);

/* Generated HIP source of a compiled kernel (NUL-terminated, owned by the
   library, valid until GmlStop); NULL if the index is invalid. */
const char *GmlxGetKernelSource(size_t GmlIdx, int KrnIdx);

/* Group size and XCD remap flag the launch-shape tuner settled on for a
   kernel (0 in both while it is still tuning). */
int    GmlxGetKernelShape(size_t GmlIdx, int KrnIdx, int *GrpSiz, int *Remap);

/* Pin the launch shape of a kernel (disables its tuner). */
int    GmlxSetKernelShape(size_t GmlIdx, int KrnIdx, int GrpSiz, int Remap);

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
   build, and return 1.  Uplinks (lower to higher dimension) are built on the
   GPU when SrcTyp is GmlVertices. */
int    GmlxBuildLink(size_t GmlIdx, int SrcTyp, int DstTyp);

/* Copy an uplink out in the reference's row-major form: deg[NmbSrc] and
   ball[NmbSrc * width] with entries (element << 4) | local index in
   descending element order, 0-padded.  width must be >= the max degree. */
int    GmlxGetUplinkTable(size_t GmlIdx, int SrcTyp, int DstTyp, int *deg, int *ball, int width);

/* --- synthetic meshes ------------------------------------------------------ */

/* Cube [0,1]^3 split in n^3 cells of 6 positively oriented (Kuhn) tets.
   crd: (n+1)^3 * 4 floats (x,y,z,0); ref: (n+1)^3 ints (1 on the boundary,
   0 inside); tet: 6 n^3 * 4 ints.  Interior vertices are jittered by a
   seeded uniform offset in [-jitter*h, jitter*h] per axis, boundary vertices
   only along their face.  order = 1 renumbers vertices and cells along a 3D
   Hilbert curve, 0 keeps lexicographic order.  Returns 1. */
int    GmlxCubeMesh(int n, double jitter, unsigned long long seed, int order,
                    float *crd, int *ref, int *tet);

#ifdef __cplusplus
}
#endif

#endif
