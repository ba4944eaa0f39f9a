/*
 * A GMlib user program written against the drop-in header only: build a
 * small Kuhn cube mesh, compute every tet's mean-ratio quality with a user
 * body through GmlCompileKernel / GmlLaunchKernel, then the minimum and the
 * mean with GmlReduceVector, and read one line back with GmlGetDataLine.
 *
 *   gcc -I include mesh_quality.c -L gmlib_amd -lgml -Wl,-rpath,gmlib_amd -o mesh_quality
 *   ./mesh_quality [cells per side]
 */
#include <stdio.h>
#include <stdlib.h>
#include "gmlib3.h"
#include "gmlx.h"

static const char *quality =
   "{\n"
   "   float4 e0 = VerCrd[1] - VerCrd[0], e1 = VerCrd[2] - VerCrd[0], e2 = VerCrd[3] - VerCrd[0];\n"
   "   float4 e3 = VerCrd[2] - VerCrd[1], e4 = VerCrd[3] - VerCrd[1], e5 = VerCrd[3] - VerCrd[2];\n"
   "   float  s  = dot(e0, e0) + dot(e1, e1) + dot(e2, e2) + dot(e3, e3) + dot(e4, e4) + dot(e5, e5);\n"
   "   float  v  = dot(cross(e0, e1), e2) / 6.;\n"
   "   qal = 12. * cbrt(9. * v * v) / s;\n"
   "}\n";

static const char *params = "typedef struct { int foo; float res; } GmlParSct;";

int main(int argc, char **argv)
{
   int n = argc > 1 ? atoi(argv[1]) : 8;
   int nv = (n + 1) * (n + 1) * (n + 1), nt = 6 * n * n * n, VerIdx, TetIdx, QalIdx, krn, cnt, idx;
   float *crd = malloc(sizeof(float) * 4 * nv), qal;
   int *ref = malloc(sizeof(int) * nv), *tet = malloc(sizeof(int) * 4 * nt), tref = 0;
   double qmin, qsum;
   size_t gml;

   GmlxCubeMesh(n, 0.2, 1, 1, crd, ref, tet);
   if(!(gml = GmlInit(0)))
      return 1;
   GmlNewParameters(gml, 8, (char *)params);
   VerIdx = GmlNewMeshData(gml, GmlVertices, nv);
   TetIdx = GmlNewMeshData(gml, GmlTetrahedra, nt);
   GmlSetDataBlock(gml, GmlVertices, 0, nv - 1, crd, crd + 4 * (nv - 1), ref, ref + nv - 1);
   GmlSetDataBlock(gml, GmlTetrahedra, 0, nt - 1, tet, tet + 4 * (nt - 1), &tref, &tref);
   QalIdx = GmlNewSolutionData(gml, GmlTetrahedra, 1, GmlFlt, "qal");
   krn = GmlCompileKernel(gml, (char *)quality, "quality", GmlTetrahedra, 2,
                          VerIdx, GmlReadMode, 0, QalIdx, GmlWriteMode, 0);
   if(!krn || GmlLaunchKernel(gml, krn) < 0)
      return 2;
   if(GmlReduceVector(gml, QalIdx, GmlMin, &qmin) < 0 || GmlReduceVector(gml, QalIdx, GmlSum, &qsum) < 0)
      return 3;
   if(!GmlGetMeshInfo(gml, GmlTetrahedra, &cnt, &idx) || idx != TetIdx || cnt != nt)
      return 4;
   GmlGetDataLine(gml, QalIdx, 0, &qal);
   printf("%d tets: min quality %.7f, mean %.7f, tet 0 %.7f, kernel %.3g s\n",
          nt, qmin, qsum / nt, qal, GmlGetKernelRunTime(gml, krn));
   GmlFreeData(gml, QalIdx);
   GmlStop(gml);
   free(crd); free(ref); free(tet);
   return 0;
}
