// Tile plan of the fused smoothing sweep (host-side construction).
//
// Vertices are cut into tiles of TV consecutive (Hilbert-ordered) vertices,
// one work-group each.  For a tile the plan lists
//   - its distinct ball tets, ascending (the union of the tiles' ball rows),
//     each as the 4 tile-local indices of its vertices in node order,
//   - the distinct vertices of those tets, ascending (global indices),
//   - for every ball slot of every vertex (SELL-64 layout of the ball), the
//     tile-local index of the slot's tet (uint16).
// The kernel stages the tile's vertex coordinates in LDS with coalesced
// loads of the ascending vertex list, forms every distinct barycentre once in
// LDS, and sums each vertex's ball from LDS in the ball's own order -- the
// same additions, in the same order, as the scatter + gather kernels.
//
// The operator form (op = 1) replaces the tet lists by each vertex's 1-ring:
// the mean ball barycentre of v is sum_u c_vu x_u / (4 deg v), c_vu =
// number of ball tets of v that contain u (c_vv = deg v), so a sweep applies
// a fixed sparse operator.  Its tile-local numbering puts the tile's own
// vertices first (local index v - v0, staged with coalesced loads) and the
// halo (the other 1-ring vertices, ascending) from TV on; only the halo is
// listed.  The ring of v is stored as 16-bit entries (tile-local index << 4
// | c_vu), ascending local index (a count above 15 continues in the next
// entry), SELL-64, zero past a ring's end.  Each ball tet of v contributes
// its three other vertex positions, so the counts sum to 3 deg v and the
// kernel derives deg v from them; v itself appears only when a ball tet
// lists it more than once (the count beyond c_vv = deg v).

#include "gml_internal.h"

#include <algorithm>
#include <atomic>
#include <thread>

namespace gml {

SweepPlan *NewSweepPlan(GmlCtx *g, int TV, int op)
{
   Uplink *u = &g->up[GmlVertices][GmlTetrahedra];
   const long long nv = g->NmbEle[GmlVertices];
   const int ti = g->TypIdx[GmlTetrahedra];
   if(!u->BalIdx || !ti || TV < 64 || (TV & 63) || TV > 1024) return nullptr;
   Data *td = &g->dat[ti];
   if(op)
   {
      // the operator form on the GPU for tiles of up to 256 vertices; the
      // host below takes larger tiles and the tiles the GPU builder declines
      SweepPlan *p = new SweepPlan();
      long long sum_halo = 0;
      int max_halo = 0;
      const int rc = GpuSweepOpPlan(g, (const int *)g->dat[u->BalIdx].dev, (const unsigned int *)g->dat[u->OffIdx].dev,
                                    (const int *)g->dat[u->DegIdx].dev, (const int *)td->dev, nv, TV, &p->vofs, &p->vlist,
                                    &p->noff, &p->nbr, &p->nslab, &sum_halo, &max_halo, &p->sum_ring);
      if(rc == 1)
      {
         p->TV = TV;
         p->ntile = (nv + TV - 1) / TV;
         p->max_nv = max_halo + TV;
         p->sum_nv = sum_halo + nv;
         p->bytes = (size_t)(p->ntile + 1) * 4 + (size_t)sum_halo * 4 + 4 + (size_t)((nv + 63) / 64 + 1) * 4
                  + (size_t)p->nslab * 128 + 128;
         g->MemSiz += p->bytes;
         DbgPrint(g, "Operator sweep plan (GPU): %lld tiles of %d vertices, %lld halo vertices, %lld ring slabs, %zu MB\n",
                  p->ntile, TV, sum_halo, p->nslab, p->bytes >> 20);
         return p;
      }
      delete p;
      if(rc < 0) return nullptr;
   }
   if(!td->host_valid && !Download(g, ti)) return nullptr;
   const int *tet = (const int *)td->host;
   const long long nchunk = (nv + 63) / 64;
   std::vector<int> deg((size_t)nv);
   std::vector<unsigned int> off((size_t)nchunk + 1);
   std::vector<int> ent((size_t)u->nslab * 64 + 64);
   if(!DevCopy(g, deg.data(), g->dat[u->DegIdx].dev, (size_t)nv * 4, 2)
   || !DevCopy(g, off.data(), g->dat[u->OffIdx].dev, off.size() * 4, 2)
   || !DevCopy(g, ent.data(), g->dat[u->BalIdx].dev, ent.size() * 4, 2))
      return nullptr;

   const long long ntile = (nv + TV - 1) / TV;
   std::vector<std::vector<int>> tv((size_t)ntile);
   std::vector<std::vector<unsigned short>> tl((size_t)ntile);
   std::vector<unsigned short> lidx(op ? 1 : (size_t)u->nslab * 64 + 64, 0);
   std::vector<std::vector<unsigned short>> ring(op ? (size_t)nv : 0);
   std::atomic<bool> fits(true);

   unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
   std::vector<std::thread> th;
   for(unsigned w = 0; w < nth; w++)
      th.emplace_back([&, w]() {
         std::vector<int> ts, vs;
         for(long long t = w; t < ntile; t += nth)
         {
            const long long v0 = t * TV, v1 = std::min(nv, v0 + TV);
            ts.clear();
            for(long long v = v0; v < v1; v++)
               for(int k = 0; k < deg[v]; k++)
                  ts.push_back(ent[((size_t)off[v >> 6] + k) * 64 + (v & 63)] >> 4);
            std::sort(ts.begin(), ts.end());
            ts.erase(std::unique(ts.begin(), ts.end()), ts.end());
            vs.clear();
            for(int e : ts)
               for(int j = 0; j < 4; j++) vs.push_back(tet[(size_t)e * 4 + j]);
            std::sort(vs.begin(), vs.end());
            vs.erase(std::unique(vs.begin(), vs.end()), vs.end());
            if(ts.size() > 65535 || vs.size() > 65535) { fits = false; continue; }
            if(op)
            {
               // tile-local numbering: the tile's own vertices first (local
               // v - v0), then its halo (the other 1-ring vertices) from TV
               // on, ascending; the plan stores the halo list only
               std::vector<int> halo;
               for(int x : vs)
                  if(x < v0 || x >= v1) halo.push_back(x);
               auto loc = [&](int x) -> unsigned int {
                  if(x >= v0 && x < v1) return (unsigned int)(x - v0);
                  return (unsigned int)(TV + (std::lower_bound(halo.begin(), halo.end(), x) - halo.begin()));
               };
               // 1-ring counts of every tile vertex, ascending tile-local index
               std::vector<unsigned int> cnt;
               for(long long v = v0; v < v1; v++)
               {
                  cnt.clear();
                  for(int k = 0; k < deg[v]; k++)
                  {
                     int e = ent[((size_t)off[v >> 6] + k) * 64 + (v & 63)] >> 4;
                     for(int j = 0; j < 4; j++) cnt.push_back(loc(tet[(size_t)e * 4 + j]));
                  }
                  std::sort(cnt.begin(), cnt.end());
                  auto &R = ring[v];
                  R.clear();
                  const unsigned int self = (unsigned int)(v - v0);
                  for(size_t i = 0; i < cnt.size();)
                  {
                     size_t j = i;
                     while(j < cnt.size() && cnt[j] == cnt[i]) j++;
                     if(cnt[i] > 4095) fits = false;
                     size_t n = j - i;
                     if(cnt[i] == self) n = n > (size_t)deg[v] ? n - deg[v] : 0;
                     for(size_t c = n; c > 0; c -= std::min<size_t>(c, 15))
                        R.push_back((unsigned short)((cnt[i] << 4) | std::min<size_t>(c, 15)));
                     i = j;
                  }
                  if(R.size() > 65535) fits = false;
               }
               tv[t] = halo;
               continue;
            }
            auto &L = tl[t];
            L.resize(ts.size() * 4);
            for(size_t i = 0; i < ts.size(); i++)
               for(int j = 0; j < 4; j++)
                  L[i * 4 + j] = (unsigned short)(std::lower_bound(vs.begin(), vs.end(), tet[(size_t)ts[i] * 4 + j]) - vs.begin());
            for(long long v = v0; v < v1; v++)
               for(int k = 0; k < deg[v]; k++)
               {
                  size_t s = ((size_t)off[v >> 6] + k) * 64 + (v & 63);
                  lidx[s] = (unsigned short)(std::lower_bound(ts.begin(), ts.end(), ent[s] >> 4) - ts.begin());
               }
            tv[t] = vs;
         }
      });
   for(auto &x : th) x.join();
   if(!fits) return nullptr;

   SweepPlan *p = new SweepPlan();
   p->TV = TV;
   p->ntile = ntile;
   std::vector<unsigned int> vofs((size_t)ntile + 1, 0), tofs((size_t)ntile + 1, 0);
   for(long long t = 0; t < ntile; t++)
   {
      vofs[t + 1] = vofs[t] + (unsigned int)tv[t].size();
      tofs[t + 1] = tofs[t] + (unsigned int)(tl[t].size() / 4);
      p->max_nv = std::max(p->max_nv, (int)tv[t].size());
      p->max_nt = std::max(p->max_nt, (int)(tl[t].size() / 4));
   }
   std::vector<int> vlist(vofs[ntile] + 1);
   if(op)
   {
      // staged per tile: TV own slots + the halo
      p->max_nv += TV;
      for(long long t = 0; t < ntile; t++)
         std::copy(tv[t].begin(), tv[t].end(), vlist.begin() + vofs[t]);
      std::vector<unsigned int> noff((size_t)nchunk + 1, 0);
      for(long long c = 0; c < nchunk; c++)
      {
         unsigned int w = 0;
         for(long long v = c * 64; v < std::min(nv, c * 64 + 64); v++) w = std::max(w, (unsigned int)ring[v].size());
         noff[c + 1] = noff[c] + w;
      }
      p->nslab = noff[nchunk];
      std::vector<unsigned short> nbr((size_t)p->nslab * 64 + 64, 0);
      for(long long v = 0; v < nv; v++)
      {
         p->sum_ring += (long long)ring[v].size();
         for(size_t k = 0; k < ring[v].size(); k++)
            nbr[((size_t)noff[v >> 6] + k) * 64 + (v & 63)] = ring[v][k];
      }
      p->vofs = (unsigned int *)DevAlloc(g, vofs.size() * 4);
      p->vlist = (int *)DevAlloc(g, vlist.size() * 4);
      p->noff = (unsigned int *)DevAlloc(g, noff.size() * 4);
      p->nbr = (unsigned short *)DevAlloc(g, nbr.size() * 2);
      if(!p->vofs || !p->vlist || !p->noff || !p->nbr
      || !DevCopy(g, p->vofs, vofs.data(), vofs.size() * 4, 1)
      || !DevCopy(g, p->vlist, vlist.data(), vlist.size() * 4, 1)
      || !DevCopy(g, p->noff, noff.data(), noff.size() * 4, 1)
      || !DevCopy(g, p->nbr, nbr.data(), nbr.size() * 2, 1))
      {
         FreeSweepPlan(g, p);
         return nullptr;
      }
      p->bytes = vofs.size() * 4 + vlist.size() * 4 + noff.size() * 4 + nbr.size() * 2;
      p->sum_nv = vofs[ntile] + nv;
      g->MemSiz += p->bytes;
      DbgPrint(g, "Operator sweep plan: %lld tiles of %d vertices, %u tile vertices, %lld ring slabs, %zu MB\n",
               ntile, TV, vofs[ntile], p->nslab, p->bytes >> 20);
      return p;
   }
   std::vector<unsigned short> tloc((size_t)tofs[ntile] * 4 + 4);
   for(long long t = 0; t < ntile; t++)
   {
      std::copy(tv[t].begin(), tv[t].end(), vlist.begin() + vofs[t]);
      std::copy(tl[t].begin(), tl[t].end(), tloc.begin() + (size_t)tofs[t] * 4);
   }
   p->vofs = (unsigned int *)DevAlloc(g, vofs.size() * 4);
   p->tofs = (unsigned int *)DevAlloc(g, tofs.size() * 4);
   p->vlist = (int *)DevAlloc(g, vlist.size() * 4);
   p->tloc = (unsigned short *)DevAlloc(g, tloc.size() * 2);
   p->lidx = (unsigned short *)DevAlloc(g, lidx.size() * 2);
   if(!p->vofs || !p->tofs || !p->vlist || !p->tloc || !p->lidx
   || !DevCopy(g, p->vofs, vofs.data(), vofs.size() * 4, 1)
   || !DevCopy(g, p->tofs, tofs.data(), tofs.size() * 4, 1)
   || !DevCopy(g, p->vlist, vlist.data(), vlist.size() * 4, 1)
   || !DevCopy(g, p->tloc, tloc.data(), tloc.size() * 2, 1)
   || !DevCopy(g, p->lidx, lidx.data(), lidx.size() * 2, 1))
   {
      FreeSweepPlan(g, p);
      return nullptr;
   }
   p->bytes = vofs.size() * 4 + tofs.size() * 4 + vlist.size() * 4 + tloc.size() * 2 + lidx.size() * 2;
   p->sum_nv = vofs[ntile];
   p->sum_nt = tofs[ntile];
   g->MemSiz += p->bytes;
   DbgPrint(g, "Sweep plan: %lld tiles of %d vertices, %u distinct tets (%.2f per tet), %u tile vertices, "
            "max %d tets / %d vertices per tile, %zu MB\n", ntile, TV, tofs[ntile],
            (double)tofs[ntile] / (double)g->NmbEle[GmlTetrahedra], vofs[ntile], p->max_nt, p->max_nv, p->bytes >> 20);
   return p;
}

StagePlan *NewStagePlan(GmlCtx *g, int Src, int Dst, int grp, long long nrows)
{
   Uplink *u = &g->up[Src][Dst];
   if(!u->BalIdx || grp < 64 || grp > 256 || (grp & 63) || nrows < 1 || nrows > g->NmbEle[Src]) return nullptr;
   StagePlan *p = new StagePlan();
   p->grp = grp;
   p->nrows = nrows;
   p->ntile = (nrows + grp - 1) / grp;
   const size_t nslot = (size_t)u->nslab * 64 + 64;
   p->lidx = (unsigned short *)DevAlloc(g, nslot * 2);
   long long sum = 0;
   int maxn = 0;
   if(!p->lidx
   || GpuStagePlan(g, (const int *)g->dat[u->BalIdx].dev, (const unsigned int *)g->dat[u->OffIdx].dev,
                   (const int *)g->dat[u->DegIdx].dev, nrows, grp, nslot, &p->tofs, &p->tlist, p->lidx, &sum, &maxn) != 1)
   {
      DevFree(g, p->lidx);
      delete p;
      return nullptr;
   }
   p->max_nt = maxn;
   p->bytes = (size_t)(p->ntile + 1) * 4 + (size_t)sum * 4 + 4 + nslot * 2;
   g->MemSiz += p->bytes;
   DbgPrint(g, "Staging plan %d->%d: %lld tiles of %d rows, %lld tile entities (%.2f per entity), max %d per tile\n",
            Src, Dst, p->ntile, grp, sum, g->NmbEle[Dst] ? (double)sum / (double)g->NmbEle[Dst] : 0., p->max_nt);
   return p;
}

void FreeStagePlan(GmlCtx *g, StagePlan *p)
{
   if(!p) return;
   DevFree(g, p->tofs); DevFree(g, p->tlist); DevFree(g, p->lidx);
   g->MemSiz -= p->bytes;
   delete p;
}

void FreeSweepPlan(GmlCtx *g, SweepPlan *p)
{
   if(!p) return;
   DevFree(g, p->vofs); DevFree(g, p->tofs); DevFree(g, p->vlist);
   DevFree(g, p->tloc); DevFree(g, p->lidx);
   DevFree(g, p->nbr); DevFree(g, p->noff);
   g->MemSiz -= p->bytes;
   delete p;
}

} // namespace gml

namespace gml {

// ---------------------------------------------------------------------------
// Element -> vertex tile plan (DownPlan): the element table of kind typ in
// tiles of kDownTile rows; per tile its distinct vertex ids, ascending, and
// per row the tile-local index of each slot, one byte each.  Built on the
// host from the table's device contents with up to 16 threads.  A table
// whose tiles reach more than 256 distinct vertices (a poorly ordered mesh:
// the tile would share little anyway) or that holds an id outside the
// vertex table gets no plan.
// ---------------------------------------------------------------------------
// per tile of TR rows: the distinct ids and the rows' byte indices; false
// if a tile exceeds 256 ids or holds an id outside [0, nv)
static DownPlan *BuildDownPlan(GmlCtx *g, int typ)
{
   const int idx = g->TypIdx[typ];
   Data *d = &g->dat[idx];
   DownPlan *p = new DownPlan();
   p->typ = typ;
   p->w = d->NmbItm * d->ItmLen;
   p->wq = (p->w + 3) / 4;
   p->nrows = d->NmbLin;
   p->ver = d->ver;
   p->src = d->dev;
   const long long nv = g->NmbEle[GmlVertices], n = p->nrows, w = p->w;
   if(w < 1 || w > 16 || n < 1 || !nv) { p->none = true; return p; }
   // the largest tile of at most 512 rows whose tiles each reach at most 256
   // vertices (GpuTilePlan, one work-group per tile)
   int TR = kDownTile, rc = 0;
   long long sum = 0;
   int maxn = 0;
   while(TR * w > 8192 && TR > 64) TR /= 2;
   for(; TR >= 64; TR /= 2)
   {
      rc = GpuTilePlan(g, (const int *)d->dev, n, (int)w, p->wq, nv, TR, &p->tofs, &p->tlist, &p->lid, &sum, &maxn);
      if(rc != 0) break;
   }
   if(rc != 1) { p->none = true; return p; }
   p->rows = TR;
   p->ntile = (n + TR - 1) / TR;
   p->sum_nv = sum;
   p->max_nv = maxn;
   p->bytes = (size_t)(p->ntile + 1) * 4 + (size_t)sum * 4 + 4 + (size_t)n * p->wq * 4;
   g->MemSiz += p->bytes;
   DbgPrint(g, "Element tile plan of %s: %lld tiles of %d rows, %lld tile vertices (%.2f per vertex), max %d per tile\n",
            BalTypStr[typ], p->ntile, TR, p->sum_nv, (double)p->sum_nv / (double)nv, p->max_nv);
   return p;
}

DownPlan *GetDownPlan(GmlCtx *g, int typ)
{
   if(g->host_only || typ <= GmlVertices || typ >= GmlMaxEleTyp || !g->TypIdx[typ] || !g->TypIdx[GmlVertices])
      return nullptr;
   Data *d = &g->dat[g->TypIdx[typ]];
   DownPlan *&p = g->dplan[typ];
   if(p && (p->ver != d->ver || p->src != d->dev || p->nrows != d->NmbLin)) { FreeDownPlan(g, p); p = nullptr; }
   if(!p) p = BuildDownPlan(g, typ);
   return p->none ? nullptr : p;
}

DownPlan *RepeatDownPlan(GmlCtx *g, int typ)
{
   if(g->host_only || typ <= GmlVertices || typ >= GmlMaxEleTyp || !g->TypIdx[typ] || !g->TypIdx[GmlVertices])
      return nullptr;
   const Data *d = &g->dat[g->TypIdx[typ]];
   const DownPlan *p = g->dplan[typ];
   if(p && p->ver == d->ver && p->src == d->dev && p->nrows == d->NmbLin) return p->none ? nullptr : g->dplan[typ];
   if(g->dplan_seen[typ] != d->ver)
   {
      g->dplan_seen[typ] = d->ver;
      return nullptr;
   }
   return GetDownPlan(g, typ);
}

} // namespace gml

using namespace gml;

extern "C" int GmlxGetTilePlanInfo(size_t GmlIdx, int typ, long long *info)
{
   GML_CTX(g, GmlIdx);
   if(!info || typ <= GmlVertices || typ >= GmlMaxEleTyp) return 0;
   DownPlan *p = GetDownPlan(g, typ);
   if(!p) return 0;
   info[0] = p->rows; info[1] = p->ntile; info[2] = p->sum_nv; info[3] = p->max_nv;
   info[4] = p->wq; info[5] = (long long)p->bytes;
   return 1;
}

namespace gml {

void FreeDownPlan(GmlCtx *g, DownPlan *p)
{
   if(!p) return;
   DevFree(g, p->tofs); DevFree(g, p->tlist); DevFree(g, p->lid);
   g->MemSiz -= p->bytes;
   delete p;
}

} // namespace gml
