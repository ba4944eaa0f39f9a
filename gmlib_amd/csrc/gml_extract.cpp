// Host-side mesh extraction entry points of the GMlib C API.
//
// GmlExtractEdges (sources/gmlib3.c:3719-3873) and GmlExtractFaces
// (sources/gmlib3.c:3880-4171): every distinct edge (face) of the mesh once,
// as a new GmlEdges (GmlTriangles / GmlQuadrilaterals) table.  Items of a
// table that already exists come first, in their order, with their refs;
// then the items of the element kinds above them, kind by kind, element by
// element, local item by local item (ItmEdgNod / ItmFacNod order), each at
// its first appearance, with ref 0.  As in the reference, the nodes of an
// extracted item are stored sorted ascending (its hash key order), and the
// return value is the number of items (triangles + quads for faces).
// Links built on a table that is replaced would go stale, so extraction is
// refused while such a link exists.  Device handles extract on the GPU
// (gml_items.hip: radix sorts of packed node keys); host-only handles
// use the ordered-map form below.

#include "gml_internal.h"

#include <algorithm>
#include <array>
#include <cstring>
#include <map>
#include <vector>

using namespace gml;

namespace {

typedef std::array<int, 4> Key;   // sorted nodes, unused slots -1

struct Extracted {
   std::vector<Key> items;
   std::vector<int> refs;
   std::map<Key, int> seen;
   void Add(const Key &k, int ref)
   {
      if(seen.emplace(k, (int)items.size()).second)
      {
         items.push_back(k);
         refs.push_back(ref);
      }
   }
};

Key Sorted(const int *n, int cnt)
{
   Key k = {-1, -1, -1, -1};
   for(int i = 0; i < cnt; i++) k[i] = n[i];
   std::sort(k.begin(), k.begin() + cnt);
   return k;
}

bool HasLinks(GmlCtx *g, int typ)
{
   for(int t = 0; t < GmlMaxEleTyp; t++)
      if(g->up[typ][t].BalIdx || g->up[t][typ].BalIdx || (t != GmlVertices && (g->LnkMat[typ][t] || g->LnkMat[t][typ])))
         return true;
   return false;
}

// host copy of a mesh table (nodes) and of its refs
bool HostTable(GmlCtx *g, int typ, const int **nod, const int **ref, int *len)
{
   const int idx = g->TypIdx[typ], rdx = g->RefIdx[typ];
   if(!idx) return false;
   Data *d = &g->dat[idx], *r = &g->dat[rdx];
   if((!d->host_valid && !Download(g, idx)) || (rdx && !r->host_valid && !Download(g, rdx))) return false;
   *nod = (const int *)d->host;
   *ref = rdx ? (const int *)r->host : nullptr;
   *len = d->ItmLen;
   return true;
}

// replace the table of kind typ (nodes per item nn) by the extracted items
int Install(size_t GmlIdx, GmlCtx *g, int typ, int nn, const Extracted &x)
{
   const int n = (int)x.items.size();
   if(g->TypIdx[typ])
   {
      const int old = g->TypIdx[typ], oref = g->RefIdx[typ];
      GmlFreeData(GmlIdx, old);
      if(oref && g->dat[oref].use) FreeDataSlot(g, oref);
      g->RefIdx[typ] = 0;
   }
   if(!n) return 0;
   const int idx = GmlNewMeshData(GmlIdx, typ, n);
   if(!idx) return -1;
   Data *d = &g->dat[idx], *r = &g->dat[g->RefIdx[typ]];
   if(!EnsureHost(d) || !EnsureHost(r)) return -1;
   int *t = (int *)d->host;
   const int w = d->ItmLen;
   memset(t, 0, d->MemSiz);
   for(int i = 0; i < n; i++)
      for(int j = 0; j < nn; j++) t[(size_t)i * w + j] = x.items[i][j];
   memcpy(r->host, x.refs.data(), (size_t)n * 4);
   if(!Upload(g, idx) || !Upload(g, g->RefIdx[typ])) return -1;
   return n;
}

} // namespace

extern "C" {

int GmlExtractEdges(size_t GmlIdx)
{
   GML_CTX(g, GmlIdx);
   if(HasLinks(g, GmlEdges))
   {
      fprintf(stderr, "gml: GmlExtractEdges: links built on the edge table exist; extract before building links\n");
      return 0;
   }
   if(!g->host_only)
   {
      const int n = GpuExtractItems(GmlIdx, g, GmlEdges, 2);
      if(n < 0) return 0;
      DbgPrint(g, "Hashed, setup and transfered %d edges to the GMlib.\n", n);
      return n;
   }
   Extracted x;
   const int *nod, *ref;
   int len;
   // the existing edges first, with their refs
   if(g->TypIdx[GmlEdges] && HostTable(g, GmlEdges, &nod, &ref, &len))
      for(long long i = 0; i < g->NmbEle[GmlEdges]; i++)
         x.Add(Sorted(&nod[i * len], 2), ref ? ref[i] : 0);
   for(int typ = GmlEdges + 1; typ < GmlMaxEleTyp; typ++)
   {
      if(!g->TypIdx[typ] || !HostTable(g, typ, &nod, &ref, &len)) continue;
      const int ne = NmbTpoLnk[typ][GmlEdges];
      for(long long i = 0; i < g->NmbEle[typ]; i++)
         for(int j = 0; j < ne; j++)
         {
            const int n2[2] = {nod[i * len + ItmEdgNod[typ][j][0]], nod[i * len + ItmEdgNod[typ][j][1]]};
            x.Add(Sorted(n2, 2), 0);
         }
   }
   const int n = Install(GmlIdx, g, GmlEdges, 2, x);
   if(n < 0) return 0;
   DbgPrint(g, "Hashed, setup and transfered %d edges to the GMlib.\n", n);
   return n;
}

int GmlExtractFaces(size_t GmlIdx)
{
   GML_CTX(g, GmlIdx);
   if(HasLinks(g, GmlTriangles) || HasLinks(g, GmlQuadrilaterals))
   {
      fprintf(stderr, "gml: GmlExtractFaces: links built on the face tables exist; extract before building links\n");
      return 0;
   }
   if(!g->host_only)
   {
      const int nt = GpuExtractItems(GmlIdx, g, GmlTriangles, 3);
      const int nq = nt < 0 ? -1 : GpuExtractItems(GmlIdx, g, GmlQuadrilaterals, 4);
      if(nq < 0) return 0;
      DbgPrint(g, "Hashed, setup and transfered %d triangles and %d quads to the GMlib.\n", nt, nq);
      return nt + nq;
   }
   Extracted tri, qad;
   const int *nod, *ref;
   int len;
   // existing surface triangles and quads first, with their refs
   if(g->TypIdx[GmlTriangles] && HostTable(g, GmlTriangles, &nod, &ref, &len))
      for(long long i = 0; i < g->NmbEle[GmlTriangles]; i++)
         tri.Add(Sorted(&nod[i * len], 3), ref ? ref[i] : 0);
   if(g->TypIdx[GmlQuadrilaterals] && HostTable(g, GmlQuadrilaterals, &nod, &ref, &len))
      for(long long i = 0; i < g->NmbEle[GmlQuadrilaterals]; i++)
         qad.Add(Sorted(&nod[i * len], 4), ref ? ref[i] : 0);
   // then the faces of the volume elements
   for(int typ = GmlTetrahedra; typ <= GmlHexahedra; typ++)
   {
      if(!g->TypIdx[typ] || !HostTable(g, typ, &nod, &ref, &len)) continue;
      int nf = 0;
      while(nf < 6 && ItmFacDeg[typ][nf]) nf++;
      for(long long i = 0; i < g->NmbEle[typ]; i++)
         for(int j = 0; j < nf; j++)
         {
            const int deg = ItmFacDeg[typ][j];
            int f[4];
            for(int k = 0; k < deg; k++) f[k] = nod[i * len + ItmFacNod[typ][j][k]];
            (deg == 3 ? tri : qad).Add(Sorted(f, deg), 0);
         }
   }
   const bool had_tri = g->TypIdx[GmlTriangles] || !tri.items.empty();
   const bool had_qad = g->TypIdx[GmlQuadrilaterals] || !qad.items.empty();
   int nt = 0, nq = 0;
   if(had_tri && (nt = Install(GmlIdx, g, GmlTriangles, 3, tri)) < 0) return 0;
   if(had_qad && (nq = Install(GmlIdx, g, GmlQuadrilaterals, 4, qad)) < 0) return 0;
   DbgPrint(g, "Hashed, setup and transfered %d triangles and %d quads to the GMlib.\n", nt, nq);
   return nt + nq;
}

} // extern "C"
