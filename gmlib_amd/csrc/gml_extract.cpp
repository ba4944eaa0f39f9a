// Host-side mesh extraction entry points of the GMlib C API.
//
// GmlExtractEdges (sources/gmlib3.c:3719-3873): the unique edges of every
// element kind above edges, numbered in order of first appearance (element
// index, then local edge index of ItmEdgNod), stored with sorted nodes and
// ref 0 as a new GmlEdges mesh table.  Merging with a pre-existing edge table
// is not provided.  GmlExtractFaces is outside this engine's scope.

#include "gml_internal.h"

#include <algorithm>
#include <cstring>
#include <unordered_map>

using namespace gml;

extern "C" {

int GmlExtractEdges(size_t GmlIdx)
{
   GML_CTX(g, GmlIdx);
   if(g->TypIdx[GmlEdges])
   {
      fprintf(stderr, "gml: GmlExtractEdges: an edge table already exists (merging is not provided)\n");
      return 0;
   }
   std::unordered_map<unsigned long long, int> seen;
   std::vector<int> edges;
   for(int typ = GmlEdges + 1; typ < GmlMaxEleTyp; typ++)
   {
      int idx = g->TypIdx[typ];
      if(!idx) continue;
      Data *d = &g->dat[idx];
      if(!d->host_valid && !Download(g, idx)) return 0;
      const int *e = (const int *)d->host;
      int ne = NmbTpoLnk[typ][GmlEdges];
      seen.reserve(seen.size() + (size_t)d->NmbLin * 2);
      for(long long i = 0; i < d->NmbLin; i++)
         for(int j = 0; j < ne; j++)
         {
            int a = e[i * d->ItmLen + ItmEdgNod[typ][j][0]], b = e[i * d->ItmLen + ItmEdgNod[typ][j][1]];
            if(a > b) std::swap(a, b);
            unsigned long long k = ((unsigned long long)(unsigned)a << 32) | (unsigned)b;
            if(seen.emplace(k, (int)(edges.size() / 2)).second)
            {
               edges.push_back(a);
               edges.push_back(b);
            }
         }
   }
   int n = (int)(edges.size() / 2);
   if(!n) return 0;
   int EdgIdx = GmlNewMeshData(GmlIdx, GmlEdges, n);
   if(!EdgIdx) return 0;
   Data *d = &g->dat[EdgIdx], *r = &g->dat[g->RefIdx[GmlEdges]];
   if(!EnsureHost(d) || !EnsureHost(r)) return 0;
   int *t = (int *)d->host;
   for(int i = 0; i < n; i++) { t[2 * i] = edges[2 * i]; t[2 * i + 1] = edges[2 * i + 1]; }
   memset(r->host, 0, (size_t)n * 4);
   Upload(g, EdgIdx);
   Upload(g, g->RefIdx[GmlEdges]);
   DbgPrint(g, "Hashed, setup and transfered %d edges to the GMlib.\n", n);
   return n;
}

int GmlExtractFaces(size_t GmlIdx)
{
   (void)GmlIdx;
   fprintf(stderr, "gml: GmlExtractFaces is not provided by this engine (face extraction is outside its scope)\n");
   return 0;
}

} // extern "C"
