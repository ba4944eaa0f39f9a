/*
 * gmlib3.h -- drop-in public interface of the MI355X-native GMlib kernel engine.
 *
 * Every prototype, enumerator value, flag and limit below keeps the value and
 * meaning of the reference interface (sources/gmlib3.h:46-123), so a GMlib user
 * program recompiles against this header and links libgml.so instead of
 * libGM.3 + OpenCL.  No prototype uses an OpenCL or HIP type; the OpenCL
 * include of the reference (sources/gmlib3.h:25-35) is therefore dropped.
 *
 * Implementation: C++ over the HIP runtime and hipRTC for gfx950
 * (gmlib_amd/csrc).  Extensions that have no counterpart in the reference live
 * in gmlx.h.
 */

#ifndef GMLIB3_H
#define GMLIB3_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Limits and flags: sources/gmlib3.h:46-60 */
#define GmlMaxDat    100
#define GmlMaxMat    10
#define GmlMaxVec    10
#define GmlMaxKrn    100
#define GmlMaxSrcSiz 50000
#define GmlMaxStrSiz 200
#define MaxGpu       10
#define GmlRefFlag   1
#define GmlReadMode  2
#define GmlWriteMode 4
#define GmlVoyeurs   8
#define GmlManual    16
#ifndef MAX_WORKGROUP_SIZE
#define MAX_WORKGROUP_SIZE 1024
#endif

/* Enumerations: sources/gmlib3.h:62-70 (values 0..N in declaration order) */
enum  element_type   {GmlVertices, GmlEdges, GmlTriangles, GmlQuadrilaterals,
                      GmlTetrahedra, GmlPyramids, GmlPrisms, GmlHexahedra,
                      GmlMaxEleTyp};
enum  opencl_type    {GmlInt, GmlInt2, GmlInt4, GmlInt8, GmlInt16,
                      GmlFlt, GmlFlt2, GmlFlt4, GmlFlt8, GmlFlt16,
                      GmlDbl, GmlDbl2, GmlDbl4, GmlDbl8, GmlDbl16,
                      GmlByt, GmlByt2, GmlByt4, GmlByt8, GmlByt16,
                      GmlMaxOclTyp};
enum reduction_opp   {GmlMin, GmlMax, GmlSum, GmlL0, GmlL1, GmlL2, GmlLinf, GmlMaxRed};

/* Entry points: sources/gmlib3.h:77-118, same order and signatures. */
size_t   GmlInit              (int);                                   /* gmlib3.c:385  */
void     GmlStop              (size_t);                                /* gmlib3.c:461  */
void     GmlListGPU           ();                                      /* gmlib3.c:486  */
void    *GmlNewParameters     (size_t, int, char *);                   /* gmlib3.c:525  */
int      GmlNewMeshData       (size_t, int, int);                      /* gmlib3.c:571  */
int      GmlNewSolutionData   (size_t, int, int, int, char *);         /* gmlib3.c:646  */
int      GmlNewLinkData       (size_t, int, int, int, char *);         /* gmlib3.c:693  */
int      GmlNewMatrix         (size_t, int, int, int, void *, int *, int *, int); /* gmlib3.c:1175 */
int      GmlNewVector         (size_t, int, int, void *, int);         /* gmlib3.c:1434 */
int      GmlFreeData          (size_t, int);                           /* gmlib3.c:1812 */
int      GmlSetDataLine       (size_t, int, int, ...);                 /* gmlib3.c:1841 */
int      GmlGetDataLine       (size_t, int, int, ...);                 /* gmlib3.c:1910 */
void     GmlSetCompilerOptions(size_t, char *);                        /* gmlib3.c:2206 */
int      GmlCompileKernel     (size_t, char *, char *, int, int, ...); /* gmlib3.c:2217 */
int      GmlLaunchKernel      (size_t, int);                           /* gmlib3.c:3064 */
int      GmlReduceVector      (size_t, int, int, double *);            /* gmlib3.c:3229 */
size_t   GmlGetMemoryUsage    (size_t);                                /* gmlib3.c:3648 */
size_t   GmlGetMemoryTransfer (size_t);                                /* gmlib3.c:3659 */
float    GmlGetMemoryAccess   (size_t);                                /* gmlib3.c:3670 */
float    GmlGetFlops          (size_t);                                /* gmlib3.c:3681 */
void     GmlDebugOn           (size_t);                                /* gmlib3.c:3691 */
void     GmlDebugOff          (size_t);                                /* gmlib3.c:3697 */
int      GmlExtractEdges      (size_t);                                /* gmlib3.c:3719 */
int      GmlExtractFaces      (size_t);                                /* gmlib3.c:3880 */
int      GmlSetNeighbours     (size_t, int);                           /* gmlib3.c:4178 */
int      GmlCheckFP64         (size_t);                                /* gmlib3.c:3708 */
int      GmlGetMeshInfo       (size_t, int, int *, int *);             /* gmlib3.c:4274 */
int      GmlGetLinkInfo       (size_t, int, int, int *, int *, int *, int *); /* gmlib3.c:4307 */
int      GmlSetDataBlock      (size_t, int, int, int, void *, void *, int *, int *); /* gmlib3.c:1968 */
int      GmlGetDataBlock      (size_t, int, int, int, void *, void *, int *, int *); /* gmlib3.c:2042 */
double   GmlGetKernelRunTime  (size_t, int);                           /* gmlib3.c:4343 */
double   GmlGetReduceRunTime  (size_t, int);                           /* gmlib3.c:4373 */
double   GmlGetWallClock      ();                                      /* gmlib3.c:4391 */
int      GmlUploadParameters  (size_t);                                /* gmlib3.c:2177 */
int      GmlDownloadParameters(size_t);                                /* gmlib3.c:2192 */
float    GmlEvaluateNumbering (size_t);                                /* gmlib3.c:4409 */
void     GmlIncludeUserToolkit(size_t, char *);                        /* gmlib3.c:4449 */
int      GmlMultMatVec        (size_t, int, int, int);                 /* gmlib3.c:3339 */
int      GmlMultDiagMatVec    (size_t, int, int, int);                 /* gmlib3.c:3582 */
int      GmlAddVec3           (size_t, int, int, int, int);            /* gmlib3.c:3411 */
int      GmlScaleVec          (size_t, int, double *);                 /* gmlib3.c:3495 */
int      GmlNormVec           (size_t, int, int, double *);            /* gmlib3.c:3534 */

/* Mesh and solution files (sources/gmlib3.h:120-123, there under
   WITH_LIBMESHB): 0-terminated lists of libMeshb keyword codes (import) or
   data indices (export); the codes are in gml_gmf.h or libMeshb's header. */
int      GmlImportMesh        (size_t, char *, ...);                   /* gmlib3.c:4482 */
int      GmlExportSolution    (size_t, char *, ...);                   /* gmlib3.c:4600 */

#ifdef __cplusplus
}
#endif

#endif
