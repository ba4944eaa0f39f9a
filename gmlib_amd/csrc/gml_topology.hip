// Link builders of libgml.
//
// Uplinks (ball of a vertex, shell of an edge or face) follow the ordering
// contract of the reference's hash builder (sources/gmlib3.c:743-1168,
// AddHsh/GetHsh 1645-1721): the entries of a row are the elements containing
// the source entity in DESCENDING element index, encoded (element << 4) |
// local index of the entity in that element.  Rows up to LenMatBas wide are
// "base" rows; if some row is wider, rows from the first such entity on are
// "high" rows and get a second kernel (gmlib3.c:911-964, 2530-2577), sized
// here from the true maximum degree.
//
// Storage is SELL-64: chunk c of 64 consecutive source entities keeps its
// entries slab-major, ent[(off[c] + k) * 64 + lane], k < width(c) = the
// chunk's largest degree, so a wave reads slot k of 64 rows as one 256-byte
// coalesced load.  Vertex balls are built on the GPU (degree count, whose
// atomic returns also give each (element, node) its slot; chunk widths,
// offset scan, atomic-free fill, per-row descending sort in registers); edge
// and face shells, downlinks and neighbours of a device handle are built on
// the GPU from sorted item keys (gml_items.hip), those of a host-only
// handle by the host forms below.

#include "gml_internal.h"

#include <algorithm>
#include <array>
#include <cstring>

namespace gml {

// ---------------------------------------------------------------------------
// GPU vertex-ball builder
// ---------------------------------------------------------------------------
// degree of every vertex, and for each node j of element e a distinct slot
// rank[e * nnod + j] in its vertex's row (in no particular order: rows are
// sorted afterwards).  The block's element nodes are first aggregated in an
// LDS hash table (local counts and local ranks by LDS atomics), then one
// global atomic per distinct vertex of the block reserves its slots: with
// locality-ordered meshes a block of 256 tets touches ~10x fewer vertices
// than it has nodes.
static const int kCntBlock = 256;
static const int kCntHash = 4096;   // >= 2 x kCntBlock x 8 nodes

__global__ void __launch_bounds__(kCntBlock) k_ball_count(const int *__restrict__ ele, int stride, int nnod,
                                                          long long nele, long long nver, int *__restrict__ deg,
                                                          int *__restrict__ rank, int *__restrict__ bad)
{
   __shared__ int hkey[kCntHash], hcnt[kCntHash];   // hcnt holds the base once reserved
   for(int i = threadIdx.x; i < kCntHash; i += kCntBlock) { hkey[i] = -1; hcnt[i] = 0; }
   __syncthreads();
   const long long e = (long long)blockIdx.x * kCntBlock + threadIdx.x;
   int slot[8], loc[8];
   for(int j = 0; j < nnod; j++)
   {
      slot[j] = -1;
      if(e >= nele) continue;
      const int v = ele[e * stride + j];
      if(v < 0 || v >= nver) { atomicOr(bad, 1); continue; }
      unsigned int h = ((unsigned int)v * 2654435761u) & (kCntHash - 1);
      for(;;)
      {
         const int k = atomicCAS(&hkey[h], -1, v);
         if(k == -1 || k == v) break;
         h = (h + 1) & (kCntHash - 1);
      }
      slot[j] = (int)h;
      loc[j] = atomicAdd(&hcnt[h], 1);
   }
   __syncthreads();
   for(int i = threadIdx.x; i < kCntHash; i += kCntBlock)
      if(hkey[i] >= 0) hcnt[i] = atomicAdd(&deg[hkey[i]], hcnt[i]);
   __syncthreads();
   if(e >= nele) return;
   for(int j = 0; j < nnod; j++)
      rank[e * nnod + j] = slot[j] >= 0 ? hcnt[slot[j]] + loc[j] : -1;
}

// width of each 64-row chunk = its largest degree, and the first row wider
// than the base width (the maximum degree is the largest chunk width, taken
// by the scan)
__global__ void k_ball_chunks(const int *__restrict__ deg, long long nver, unsigned int *__restrict__ w,
                              int *__restrict__ stats, int baswid)
{
   long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   int d = v < nver ? deg[v] : 0;
   int m = d;
   for(int o = 32; o > 0; o >>= 1)
      m = max(m, __shfl_xor(m, o, 64));
   if((threadIdx.x & 63) == 0 && v < nver + 63)
   {
      long long c = v >> 6;
      if(c * 64 < nver) w[c] = (unsigned int)m;
   }
   if(v < nver && d > baswid)
      atomicMin(&stats[1], (int)v);
}

// exclusive scan of the chunk widths (one block; nchunk is at most a few
// million, each thread scans a contiguous run)
__global__ void k_ball_scan(const unsigned int *__restrict__ w, long long nchunk, unsigned int *__restrict__ off,
                            int *__restrict__ stats)
{
   __shared__ unsigned long long part[1024];
   __shared__ unsigned int wmax[1024];
   const int t = threadIdx.x, nt = blockDim.x;
   long long per = (nchunk + nt - 1) / nt;
   long long b = per * t, e = min(b + per, nchunk);
   unsigned long long s = 0;
   unsigned int mx = 0;
   for(long long i = b; i < e; i++) { s += w[i]; mx = max(mx, w[i]); }
   part[t] = s;
   wmax[t] = mx;
   __syncthreads();
   if(t == 0)
   {
      unsigned long long acc = 0;
      unsigned int m = 0;
      for(int i = 0; i < nt; i++) { unsigned long long x = part[i]; part[i] = acc; acc += x; m = max(m, wmax[i]); }
      off[nchunk] = (unsigned int)acc;
      stats[0] = (int)m;
   }
   __syncthreads();
   unsigned long long acc = part[t];
   for(long long i = b; i < e; i++) { off[i] = (unsigned int)acc; acc += w[i]; }
}

__global__ void k_ball_fill(const int *__restrict__ ele, int stride, int nnod, long long nele, long long nver,
                            const unsigned int *__restrict__ off, const int *__restrict__ rank, int *__restrict__ ent)
{
   long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if(e >= nele) return;
   for(int j = 0; j < nnod; j++)
   {
      int v = ele[e * stride + j], p = rank[e * nnod + j];
      if(v < 0 || v >= nver || p < 0) continue;
      ent[((size_t)off[v >> 6] + p) * 64 + (v & 63)] = (int)((e << 4) | j);
   }
}

// descending sort of each row (the reference's chain order): rows of up to
// 32 entries with a bitonic network on registers (every index compile-time,
// so nothing goes to scratch), longer rows in place in memory
__global__ void k_ball_sort(const int *__restrict__ deg, long long nver, const unsigned int *__restrict__ off,
                            int *__restrict__ ent)
{
   long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if(v >= nver) return;
   int d = deg[v];
   int *row = ent + (size_t)off[v >> 6] * 64 + (v & 63);
   if(d <= 32)
   {
      int r[32];
#pragma unroll
      for(int k = 0; k < 32; k++) r[k] = k < d ? row[(size_t)k * 64] : (int)0x80000000;
#pragma unroll
      for(int sz = 2; sz <= 32; sz <<= 1)
#pragma unroll
         for(int st = sz >> 1; st > 0; st >>= 1)
#pragma unroll
            for(int i = 0; i < 32; i++)
            {
               const int j = i ^ st;
               if(j > i)
               {
                  // descending overall: blocks with (i & sz) == 0 descend
                  const bool desc = (i & sz) == 0;
                  const int a = r[i], b = r[j];
                  r[i] = desc ? max(a, b) : min(a, b);
                  r[j] = desc ? min(a, b) : max(a, b);
               }
            }
#pragma unroll
      for(int k = 0; k < 32; k++)
         if(k < d) row[(size_t)k * 64] = r[k];
   }
   else
   {
      for(int k = 1; k < d; k++)
      {
         int x = row[(size_t)k * 64], i = k - 1;
         while(i >= 0 && row[(size_t)i * 64] < x) { row[(size_t)(i + 1) * 64] = row[(size_t)i * 64]; i--; }
         row[(size_t)(i + 1) * 64] = x;
      }
   }
}

// Returns 1 on success.  deg must hold nver ints (zeroed here); *off_out and
// *ent_out are new device allocations owned by the caller.
int gpu_vertex_ball(hipStream_t s, const int *ele, int stride, int nnod, long long nele,
                    long long nver, int *deg, unsigned int **off_out, int **ent_out,
                    long long *nslab_out, int *maxdeg_out, int *firsthigh_out, int baswid)
{
   long long nchunk = (nver + 63) / 64;
   unsigned int *w = nullptr, *off = nullptr;
   int *stats = nullptr, *rank = nullptr, *ent = nullptr;
   int hs[3] = {0, 0, 0};
   const int B = 256;
   unsigned int nbe = (unsigned int)((nele + B - 1) / B), nbv = (unsigned int)((nver + 63 + B - 1) / B);
   static_assert(kCntBlock == B, "k_ball_count takes one element per thread of a 256-thread block");
   if(nnod < 1 || nnod > 8) return 0;
   *off_out = nullptr; *ent_out = nullptr;

   if(hipMalloc(&w, (nchunk + 1) * sizeof(unsigned int)) != hipSuccess
   || hipMalloc(&off, (nchunk + 1) * sizeof(unsigned int)) != hipSuccess
   || hipMalloc(&stats, 4 * sizeof(int)) != hipSuccess
   || hipMalloc(&rank, (size_t)nele * nnod * sizeof(int) + 4) != hipSuccess)
      goto fail;
   HIPCHK(hipMemsetAsync(deg, 0, nver * sizeof(int), s));
   HIPCHK(hipMemsetAsync(w, 0, (nchunk + 1) * sizeof(unsigned int), s));
   {
      int init[4] = {0, (int)0x7fffffff, 0, 0};
      HIPCHK(hipMemcpyAsync(stats, init, sizeof(init), hipMemcpyHostToDevice, s));
   }
   if(nele > 0)
      hipLaunchKernelGGL(k_ball_count, dim3(nbe), dim3(B), 0, s, ele, stride, nnod, nele, nver, deg, rank, stats + 2);
   hipLaunchKernelGGL(k_ball_chunks, dim3(nbv), dim3(B), 0, s, deg, nver, w, stats, baswid);
   hipLaunchKernelGGL(k_ball_scan, dim3(1), dim3(1024), 0, s, w, nchunk, off, stats);
   {
      unsigned int nslab = 0;
      HIPCHK(hipMemcpyAsync(&nslab, off + nchunk, sizeof(unsigned int), hipMemcpyDeviceToHost, s));
      HIPCHK(hipMemcpyAsync(hs, stats, 3 * sizeof(int), hipMemcpyDeviceToHost, s));
      if(hipStreamSynchronize(s) != hipSuccess) goto fail;
      if(hs[2])
      {
         fprintf(stderr, "gml: element table references vertices out of range\n");
         goto fail;
      }
      *nslab_out = nslab;
      if(hipMalloc(&ent, ((size_t)nslab * 64 + 64) * sizeof(int)) != hipSuccess)
         goto fail;
      HIPCHK(hipMemsetAsync(ent, 0, ((size_t)nslab * 64 + 64) * sizeof(int), s));
   }
   if(nele > 0)
      hipLaunchKernelGGL(k_ball_fill, dim3(nbe), dim3(B), 0, s, ele, stride, nnod, nele, nver, off, rank, ent);
   hipLaunchKernelGGL(k_ball_sort, dim3((unsigned int)((nver + B - 1) / B)), dim3(B), 0, s, deg, nver, off, ent);
   if(hipStreamSynchronize(s) != hipSuccess) goto fail;
   (void)hipFree(w);
   (void)hipFree(stats);
   (void)hipFree(rank);
   *maxdeg_out = hs[0];
   *firsthigh_out = hs[1] == 0x7fffffff ? -1 : hs[1];
   *off_out = off;
   *ent_out = ent;
   return 1;
fail:
   (void)hipGetLastError();
   (void)hipFree(w); (void)hipFree(off); (void)hipFree(stats); (void)hipFree(rank); (void)hipFree(ent);
   return 0;
}

// SELL-64 layout of rows with the given degrees (device array of n ints):
// chunk offsets (*off_out, n/64 + 1 entries) and a zeroed entry array
// (*ent_out, nslab * 64 + 64 ints), both new device allocations owned by the
// caller; the maximum degree and the first row wider than baswid (-1: none).
int gpu_sell_layout(hipStream_t s, const int *deg, long long n, int baswid, unsigned int **off_out,
                    int **ent_out, long long *nslab_out, int *maxdeg_out, int *firsthigh_out)
{
   const long long nchunk = (n + 63) / 64;
   const int B = 256;
   unsigned int *w = nullptr, *off = nullptr;
   int *stats = nullptr, *ent = nullptr, hs[2] = {0, 0};
   unsigned int nslab = 0;
   const int init[4] = {0, (int)0x7fffffff, 0, 0};
   *off_out = nullptr;
   *ent_out = nullptr;
   if(hipMalloc(&w, (nchunk + 1) * sizeof(unsigned int)) != hipSuccess
   || hipMalloc(&off, (nchunk + 1) * sizeof(unsigned int)) != hipSuccess
   || hipMalloc(&stats, 4 * sizeof(int)) != hipSuccess
   || hipMemsetAsync(w, 0, (nchunk + 1) * sizeof(unsigned int), s) != hipSuccess
   || hipMemcpyAsync(stats, init, sizeof(init), hipMemcpyHostToDevice, s) != hipSuccess)
      goto fail;
   hipLaunchKernelGGL(k_ball_chunks, dim3((unsigned int)((n + 63 + B - 1) / B)), dim3(B), 0, s, deg, n, w, stats, baswid);
   hipLaunchKernelGGL(k_ball_scan, dim3(1), dim3(1024), 0, s, w, nchunk, off, stats);
   if(hipMemcpyAsync(&nslab, off + nchunk, sizeof(unsigned int), hipMemcpyDeviceToHost, s) != hipSuccess
   || hipMemcpyAsync(hs, stats, 2 * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess
   || hipStreamSynchronize(s) != hipSuccess
   || hipMalloc(&ent, ((size_t)nslab * 64 + 64) * sizeof(int)) != hipSuccess
   || hipMemsetAsync(ent, 0, ((size_t)nslab * 64 + 64) * sizeof(int), s) != hipSuccess)
      goto fail;
   (void)hipFree(w);
   (void)hipFree(stats);
   *off_out = off;
   *ent_out = ent;
   *nslab_out = nslab;
   *maxdeg_out = hs[0];
   *firsthigh_out = hs[1] == 0x7fffffff ? -1 : hs[1];
   return 1;
fail:
   (void)hipGetLastError();
   (void)hipFree(w); (void)hipFree(off); (void)hipFree(stats); (void)hipFree(ent);
   return 0;
}

// ---------------------------------------------------------------------------
// Host helpers
// ---------------------------------------------------------------------------
typedef std::array<int, 4> Key;

// sorted node tuple of local item j (an edge or face) of element e
static int ItemKey(int EleTyp, const int *nod, int ItmTyp, int j, Key *k)
{
   Key t = {{-1, -1, -1, -1}};
   int n = 0;
   if(ItmTyp == GmlVertices) { t[0] = nod[j]; n = 1; }
   else if(ItmTyp == GmlEdges) { t[0] = nod[ItmEdgNod[EleTyp][j][0]]; t[1] = nod[ItmEdgNod[EleTyp][j][1]]; n = 2; }
   else
   {
      n = ItmFacDeg[EleTyp][j];
      for(int i = 0; i < n; i++) t[i] = nod[ItmFacNod[EleTyp][j][i]];
   }
   std::sort(t.begin(), t.begin() + n);
   *k = t;
   return n;
}

// number of local items of kind ItmTyp in an element of kind EleTyp
static int NmbItems(int EleTyp, int ItmTyp)
{
   if(ItmTyp == GmlVertices) return EleNmbNod[EleTyp];
   if(ItmTyp == GmlEdges) return NmbTpoLnk[EleTyp][GmlEdges];
   int n = 0, want = ItmTyp == GmlTriangles ? 3 : 4;
   for(int j = 0; j < 6; j++) if(ItmFacDeg[EleTyp][j]) n = j + 1;
   (void)want;
   return n;
}

static int FaceNodes(int ItmTyp) { return ItmTyp == GmlTriangles ? 3 : 4; }

static bool HostTable(GmlCtx *g, int typ, std::vector<int> *tab, int *stride)
{
   int idx = g->TypIdx[typ];
   if(!idx) return false;
   Data *d = &g->dat[idx];
   if(!d->host_valid && !Download(g, idx)) return false;
   *stride = d->ItmLen;
   tab->assign((const int *)d->host, (const int *)d->host + (size_t)d->NmbLin * d->ItmLen);
   return true;
}

static int NewLinkSlot(GmlCtx *g, int kind, int Src, int Dst, const std::string &nam, size_t bytes)
{
   int idx = NewDatIdx(g);
   if(!idx) return 0;
   Data *d = &g->dat[idx];
   d->kind = kind;
   d->MshTyp = Src;
   d->LnkTyp = Dst;
   d->ItmTyp = GmlInt;
   d->NmbItm = 1; d->ItmSiz = 4; d->ItmLen = 1;
   d->NmbLin = (long long)(bytes / 4);
   d->LinSiz = 4;
   d->MemSiz = bytes;
   d->nam = nam;
   return idx;
}

// Register device SELL tables (already allocated) as data slots of uplink Src->Dst.
static int RegisterUplink(GmlCtx *g, int Src, int Dst, int *degdev, unsigned int *off, int *ent,
                          long long nslab, int maxdeg, int firsthigh)
{
   Uplink *u = &g->up[Src][Dst];
   std::string b = std::string(BalTypStr[Src]) + BalTypStr[Dst];
   long long nsrc = g->NmbEle[Src], nchunk = (nsrc + 63) / 64;

   int di = NewLinkSlot(g, DkLnk, Src, Dst, b + "Deg", (size_t)nsrc * 4);
   int bi = NewLinkSlot(g, DkBal, Src, Dst, b + "Bal", ((size_t)nslab * 64 + 64) * 4);
   int oi = NewLinkSlot(g, DkOff, Src, Dst, b + "BalOff", (size_t)(nchunk + 1) * 4);
   if(!di || !bi || !oi) return 0;
   g->dat[di].dev = degdev; g->dat[di].NmbLin = nsrc;
   g->dat[bi].dev = ent; g->dat[bi].VoyNam = b + "Voy"; g->dat[bi].OffIdx = oi;
   g->dat[oi].dev = off;
   g->MemSiz += g->dat[di].MemSiz + g->dat[bi].MemSiz + g->dat[oi].MemSiz;

   u->DegIdx = di; u->BalIdx = bi; u->OffIdx = oi;
   u->BasWid = LenMatBas[Src][Dst];
   u->MaxDeg = maxdeg;
   u->nslab = nslab;
   u->SrcCnt = nsrc;
   u->DstVer = g->dat[g->TypIdx[Dst]].ver;
   u->SrcVer = Src == GmlVertices ? 0 : g->dat[g->TypIdx[Src]].ver;
   if(firsthigh >= 0 && maxdeg > u->BasWid)
   {
      int w = 1;
      while(w < maxdeg) w <<= 1;
      u->HghWid = w;
      u->MaxPos = firsthigh;
   }
   else
   {
      u->HghWid = 0;
      u->MaxPos = (int)nsrc;
   }
   DbgPrint(g, "Ball generation: type %s -> %s, max degree %d, base rows %d x %d, high rows %lld x %d, %lld slabs\n",
            BalTypStr[Src], BalTypStr[Dst], maxdeg, u->MaxPos, u->BasWid,
            u->HghWid ? nsrc - u->MaxPos : 0LL, u->HghWid, nslab);
   return 1;
}

int BuildUplink(GmlCtx *g, int Src, int Dst)
{
   if(g->up[Src][Dst].BalIdx) return 1;
   int DstIdx = g->TypIdx[Dst];
   long long nsrc = g->NmbEle[Src], ndst = g->NmbEle[Dst];
   if(!DstIdx || (Src != GmlVertices && !g->TypIdx[Src]))
   {
      fprintf(stderr, "gml: uplink %s -> %s needs both mesh tables\n", BalTypStr[Src], BalTypStr[Dst]);
      return 0;
   }
   if(ndst >= (1LL << 27))
   {
      fprintf(stderr, "gml: uplink entries encode element << 4 in 32 bits: %lld elements is too many\n", ndst);
      return 0;
   }
   int *degdev = (int *)DevAlloc(g, (size_t)nsrc * 4 + 4);
   if(!degdev) return 0;

   if(Src == GmlVertices && !g->host_only)
   {
      Data *e = &g->dat[DstIdx];
      unsigned int *off = nullptr;
      int *ent = nullptr, maxdeg = 0, first = -1;
      long long nslab = 0;
      if(!gpu_vertex_ball(g->stream, (const int *)e->dev, e->ItmLen, EleNmbNod[Dst], ndst, nsrc,
                          degdev, &off, &ent, &nslab, &maxdeg, &first, LenMatBas[Src][Dst]))
      {
         DevFree(g, degdev);
         return 0;
      }
      return RegisterUplink(g, Src, Dst, degdev, off, ent, nslab, maxdeg, first);
   }
   if(!g->host_only)
   {
      // edge / face shells: item keys sorted on the GPU
      unsigned int *off = nullptr;
      int *ent = nullptr, maxdeg = 0, first = -1;
      long long nslab = 0;
      if(!GpuShell(g, Src, Dst, degdev, &off, &ent, &nslab, &maxdeg, &first))
      {
         DevFree(g, degdev);
         return 0;
      }
      return RegisterUplink(g, Src, Dst, degdev, off, ent, nslab, maxdeg, first);
   }

   // host builder (host-only handles): rows in descending element order
   std::vector<int> dtab, stab;
   int dst = 0, sst = 0;
   if(!HostTable(g, Dst, &dtab, &dst) || (Src != GmlVertices && !HostTable(g, Src, &stab, &sst)))
   {
      DevFree(g, degdev);
      return 0;
   }
   std::vector<int> deg((size_t)nsrc, 0);
   std::vector<std::vector<int>> rows((size_t)nsrc);
   if(Src == GmlVertices)
   {
      int nn = EleNmbNod[Dst];
      for(long long e = ndst - 1; e >= 0; e--)
         for(int j = 0; j < nn; j++)
         {
            int v = dtab[(size_t)e * dst + j];
            if(v >= 0 && v < nsrc) rows[v].push_back((int)((e << 4) | j));
         }
   }
   else
   {
      // (key, element, local item) sorted by key then element desc
      int ni = NmbItems(Dst, Src);
      struct Ent { Key k; int e, j; };
      std::vector<Ent> all;
      all.reserve((size_t)ndst * ni);
      for(long long e = 0; e < ndst; e++)
         for(int j = 0; j < ni; j++)
         {
            Key k;
            int n = ItemKey(Dst, &dtab[(size_t)e * dst], Src, j, &k);
            if(Src >= GmlTriangles && n != FaceNodes(Src)) continue;
            all.push_back({k, (int)e, j});
         }
      std::sort(all.begin(), all.end(), [](const Ent &a, const Ent &b) {
         return a.k != b.k ? a.k < b.k : a.e > b.e; });
      for(long long s = 0; s < nsrc; s++)
      {
         Key k = {{-1, -1, -1, -1}};
         int n = Src == GmlEdges ? 2 : FaceNodes(Src);
         for(int i = 0; i < n; i++) k[i] = stab[(size_t)s * sst + i];
         std::sort(k.begin(), k.begin() + n);
         auto lo = std::lower_bound(all.begin(), all.end(), k, [](const Ent &a, const Key &b) { return a.k < b; });
         for(auto it = lo; it != all.end() && it->k == k; ++it)
            rows[s].push_back((it->e << 4) | it->j);
      }
   }
   int maxdeg = 0, first = -1, bw = LenMatBas[Src][Dst];
   for(long long s = 0; s < nsrc; s++)
   {
      deg[s] = (int)rows[s].size();
      maxdeg = std::max(maxdeg, deg[s]);
      if(first < 0 && deg[s] > bw) first = (int)s;
   }
   // pack SELL-64
   long long nchunk = (nsrc + 63) / 64;
   std::vector<unsigned int> off((size_t)nchunk + 1, 0);
   for(long long c = 0; c < nchunk; c++)
   {
      int w = 0;
      for(long long v = c * 64; v < std::min(nsrc, c * 64 + 64); v++) w = std::max(w, deg[v]);
      off[c + 1] = off[c] + w;
   }
   long long nslab = off[nchunk];
   std::vector<int> ent(((size_t)nslab * 64 + 64), 0);
   for(long long v = 0; v < nsrc; v++)
      for(int k = 0; k < deg[v]; k++)
         ent[((size_t)off[v >> 6] + k) * 64 + (v & 63)] = rows[v][k];
   unsigned int *doff = (unsigned int *)DevAlloc(g, off.size() * 4);
   int *dent = (int *)DevAlloc(g, ent.size() * 4);
   if(!doff || !dent)
   {
      DevFree(g, degdev); DevFree(g, doff); DevFree(g, dent);
      return 0;
   }
   if(!DevCopy(g, degdev, deg.data(), (size_t)nsrc * 4, 1)
   || !DevCopy(g, doff, off.data(), off.size() * 4, 1)
   || !DevCopy(g, dent, ent.data(), ent.size() * 4, 1))
      return 0;
   return RegisterUplink(g, Src, Dst, degdev, doff, dent, nslab, maxdeg, first);
}

// Downlink Src -> Dst (Dst of lower dimension, not the vertices): for each
// local item of each Src element, the index of the Dst entity with the same
// nodes (the highest such index, as the reference's chain head), 0 if none.
int BuildDownlink(GmlCtx *g, int Src, int Dst)
{
   if(g->LnkMat[Src][Dst]) return 1;
   if(!g->TypIdx[Src] || !g->TypIdx[Dst])
   {
      fprintf(stderr, "gml: downlink %s -> %s needs both mesh tables\n", BalTypStr[Src], BalTypStr[Dst]);
      return 0;
   }
   std::string nam = std::string(BalTypStr[Src]) + BalTypStr[Dst] + "Lnk";
   if(!g->host_only)
   {
      const int idx = GpuDownlink(g, Src, Dst, nam);
      if(!idx) return 0;
      g->LnkMat[Src][Dst] = idx;
      return 1;
   }
   std::vector<int> stab, dtab;
   int sst = 0, dst = 0;
   if(!HostTable(g, Src, &stab, &sst) || !HostTable(g, Dst, &dtab, &dst))
   {
      fprintf(stderr, "gml: downlink %s -> %s needs both mesh tables\n", BalTypStr[Src], BalTypStr[Dst]);
      return 0;
   }
   long long ns = g->NmbEle[Src], nd = g->NmbEle[Dst];
   int nn = Dst == GmlEdges ? 2 : FaceNodes(Dst);
   std::vector<std::pair<Key, int>> keys;
   keys.reserve((size_t)nd);
   for(long long i = 0; i < nd; i++)
   {
      Key k = {{-1, -1, -1, -1}};
      for(int j = 0; j < nn; j++) k[j] = dtab[(size_t)i * dst + j];
      std::sort(k.begin(), k.begin() + nn);
      keys.push_back({k, (int)i});
   }
   std::sort(keys.begin(), keys.end(), [](const std::pair<Key, int> &a, const std::pair<Key, int> &b) {
      return a.first != b.first ? a.first < b.first : a.second > b.second; });

   int w = LenMatBas[Src][Dst], cnt, vec, typ;
   GetCntVec(w, &cnt, &vec, &typ);
   int rw = cnt * vec, ni = NmbItems(Src, Dst);
   int idx = NewDatIdx(g);
   if(!idx) return 0;
   Data *d = &g->dat[idx];
   d->kind = DkLnk; d->MshTyp = Src; d->LnkTyp = Dst;
   d->ItmTyp = typ; d->NmbItm = cnt; d->ItmSiz = OclTypSiz[typ]; d->ItmLen = vec;
   d->NmbLin = ns; d->LinSiz = (size_t)rw * 4; d->MemSiz = (size_t)ns * d->LinSiz;
   d->nam = nam;
   if(!AllocData(g, d) || !EnsureHost(d)) return 0;
   int *t = (int *)d->host;
   for(long long i = 0; i < ns; i++)
      for(int j = 0; j < ni && j < rw; j++)
      {
         Key k;
         int n = ItemKey(Src, &stab[(size_t)i * sst], Dst, j, &k);
         if(Dst >= GmlTriangles && n != nn) continue;
         auto lo = std::lower_bound(keys.begin(), keys.end(), k,
                                    [](const std::pair<Key, int> &a, const Key &b) { return a.first < b; });
         if(lo != keys.end() && lo->first == k)
            t[(size_t)i * rw + j] = lo->second;
      }
   Upload(g, idx);
   g->LnkMat[Src][Dst] = idx;
   return 1;
}

void DropUplink(GmlCtx *g, int Src, int Dst)
{
   Uplink *u = &g->up[Src][Dst];
   for(StagePlan *&p : u->stage)
   {
      FreeStagePlan(g, p);
      p = nullptr;
   }
   for(int idx : {u->DegIdx, u->BalIdx, u->OffIdx})
      if(idx && g->dat[idx].use) FreeDataSlot(g, idx);
   if(Src == GmlVertices && Dst == GmlTetrahedra)
   {
      // the smoothing sweep plans are built on the vertex -> tet ball
      FreeSweepPlan(g, g->plan);
      FreeSweepPlan(g, g->oplan);
      g->plan = g->oplan = nullptr;
   }
   *u = Uplink();
}

Uplink *CurrentUplink(GmlCtx *g, int Src, int Dst)
{
   Uplink *u = &g->up[Src][Dst];
   const int di = g->TypIdx[Dst], si = Src == GmlVertices ? 0 : g->TypIdx[Src];
   if(!di || (Src != GmlVertices && !si)) return nullptr;
   if(u->BalIdx
   && (u->SrcCnt != g->NmbEle[Src] || u->DstVer != g->dat[di].ver || (si && u->SrcVer != g->dat[si].ver)))
   {
      if(u->pinned) return nullptr;
      DropUplink(g, Src, Dst);
   }
   if(!u->BalIdx && !BuildUplink(g, Src, Dst)) return nullptr;
   return u;
}

} // namespace gml

using namespace gml;

extern "C" {

int GmlxBuildLink(size_t GmlIdx, int Src, int Dst)
{
   GML_CTX(g, GmlIdx);
   if(Src < 0 || Src >= GmlMaxEleTyp || Dst < 0 || Dst >= GmlMaxEleTyp) return 0;
   if(MshTypDim[Src] < MshTypDim[Dst])
   {
      if(!BuildUplink(g, Src, Dst)) return 0;
      g->up[Src][Dst].pinned = true;
      return 1;
   }
   if(MshTypDim[Src] > MshTypDim[Dst]) return Dst == GmlVertices ? (g->TypIdx[Src] ? 1 : 0) : BuildDownlink(g, Src, Dst);
   return 0;
}

int GmlxGetLinkData(size_t GmlIdx, int Src, int Dst)
{
   GML_CTX(g, GmlIdx);
   if(Src < 0 || Src >= GmlMaxEleTyp || Dst <= GmlVertices || Dst >= GmlMaxEleTyp) return 0;
   return g->LnkMat[Src][Dst];
}

int GmlxGetUplinkTable(size_t GmlIdx, int Src, int Dst, int *deg, int *ball, int width)
{
   GML_CTX(g, GmlIdx);
   if(Src < 0 || Src >= GmlMaxEleTyp || Dst < 0 || Dst >= GmlMaxEleTyp) return 0;
   Uplink *u = &g->up[Src][Dst];
   if(!u->BalIdx || width < u->MaxDeg || u->SrcCnt != g->NmbEle[Src]) return 0;
   long long n = g->NmbEle[Src], nchunk = (n + 63) / 64;
   std::vector<unsigned int> off((size_t)nchunk + 1);
   std::vector<int> ent((size_t)u->nslab * 64 + 64);
   if(!DevCopy(g, deg, g->dat[u->DegIdx].dev, (size_t)n * 4, 2)
   || !DevCopy(g, off.data(), g->dat[u->OffIdx].dev, off.size() * 4, 2)
   || !DevCopy(g, ent.data(), g->dat[u->BalIdx].dev, ent.size() * 4, 2))
      return 0;
   memset(ball, 0, (size_t)n * width * sizeof(int));
   for(long long v = 0; v < n; v++)
      for(int k = 0; k < deg[v]; k++)
         ball[(size_t)v * width + k] = ent[((size_t)off[v >> 6] + k) * 64 + (v & 63)];
   return 1;
}

// Neighbours of the elements of kind typ (GmlSetNeighbours,
// sources/gmlib3.c:4178-4267): a link data "ngb" with, for each element and
// each of its items of kind NgbTyp[typ] (edges: vertices; triangles and
// quads: edges; tets: triangles; hexes: quads), the element sharing that
// item, 0 if none.  When more than one other element shares it, the lowest
// index is kept: the last match of the reference's chain walk.  The
// reference hands the table to GmlSetDataLine as a pointer where ints are
// read (so its entries are not the neighbours) and sizes it with the number
// of faces, too narrow for triangles and quads; here each line holds one
// entry per item (2, 3, 4, 4, 6 for edges, triangles, quads, tets, hexes,
// padded to the link vector width).  Pyramids and prisms are refused as in
// the reference.  Returns the link data index, 0 on failure.
int GmlSetNeighbours(size_t GmlIdx, int typ)
{
   static const int NgbTyp[GmlMaxEleTyp] = {-1, GmlVertices, GmlEdges, GmlEdges, GmlTriangles, GmlTriangles,
                                            GmlTriangles, GmlQuadrilaterals};
   GML_CTX(g, GmlIdx);
   if(typ == GmlPyramids || typ == GmlPrisms)
   {
      puts("Neighbours calculation between prisms or pyramids is not yet implemented.");
      return 0;
   }
   if(typ <= GmlVertices || typ >= GmlMaxEleTyp || !g->TypIdx[typ]) return 0;
   if(g->LnkMat[typ][typ])
   {
      fprintf(stderr, "gml: GmlSetNeighbours: the %s neighbours are already built\n", BalTypStr[typ]);
      return 0;
   }
   const int ngb = NgbTyp[typ], w = NmbTpoLnk[typ][ngb];
   const int idx = GmlNewLinkData(GmlIdx, typ, typ, w, (char *)"ngb");
   if(!idx) return 0;
   Data *d = &g->dat[idx];
   const int rw = d->NmbItm * d->ItmLen;
   if(!g->host_only)
   {
      if(!GpuNeighbours(g, typ, ngb, idx, rw))
      {
         g->LnkMat[typ][typ] = 0;
         FreeDataSlot(g, idx);
         return 0;
      }
      return idx;
   }
   // host-only handle: (sorted nodes, element, local item) sorted, runs of equal nodes
   std::vector<int> tab;
   int st = 0;
   if(!HostTable(g, typ, &tab, &st)) return 0;
   const long long ne = g->NmbEle[typ];
   const int ni = NmbItems(typ, ngb), nn = ngb == GmlVertices ? 1 : ngb == GmlEdges ? 2 : FaceNodes(ngb);
   struct Ent { Key k; int e, j; };
   std::vector<Ent> all;
   all.reserve((size_t)ne * ni);
   for(long long e = 0; e < ne; e++)
      for(int j = 0; j < ni; j++)
      {
         Key k;
         if(ItemKey(typ, &tab[(size_t)e * st], ngb, j, &k) != nn) continue;
         all.push_back({k, (int)e, j});
      }
   std::sort(all.begin(), all.end(), [](const Ent &a, const Ent &b) {
      return a.k != b.k ? a.k < b.k : (a.e != b.e ? a.e < b.e : a.j < b.j); });
   int *t = (int *)d->dev;
   for(size_t a = 0, b; a < all.size(); a = b)
   {
      for(b = a; b < all.size() && all[b].k == all[a].k; b++) {}
      for(size_t i = a; i < b; i++)
      {
         int r = 0;
         for(size_t k = a; k < b; k++)
            if(all[k].e != all[i].e) { r = all[k].e; break; }
         if(all[i].j < rw) t[(size_t)all[i].e * rw + all[i].j] = r;
      }
   }
   d->host_valid = false;
   BumpVersion(d);
   return idx;
}

} // extern "C"
