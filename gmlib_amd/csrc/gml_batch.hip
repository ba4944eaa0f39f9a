// Batched mesh quality: many independent tet meshes kept on the device and
// processed together by one persistent kernel, replayed from a HIP graph.
//
// Work per run (the MeshQuality + BallQuality patterns of the reference,
// examples/MeshQuality and examples/BallQuality, applied to every mesh of
// the batch at once):
//   - the mean-ratio quality of every tet (the expression of
//     tet_meanratio.cl; coordinates are stored with a zero fourth component,
//     so meanratio3 gives GmlxTetQuality's values bit for bit),
//   - the minimum of it over every vertex ball (as GmlxBallMin: an empty
//     ball gives 0),
//   - per mesh, the minimum and the sum of the tet qualities (FP64 sum).
//
// Layout.  Each mesh is cut into tiles of TV consecutive vertices (tiles
// never straddle two meshes).  A tet belongs to the tile that owns its node
// 0, and tets are numbered tile by tile in that order (the plan keeps the
// permutation back to the user's order), so a tile's own tets are one
// contiguous range.  A tile lists its distinct ball tets -- its own first,
// then the halo -- as 4 tile-local vertex indices (uint16), the distinct
// vertices of those tets (global indices, ascending), and, for each of its
// own vertices, its ball as tile-local tet indices in SELL-64 slabs whose
// rows are padded to a multiple of Pad with a sentinel that points at +inf.
// One work-group processes a tile: stage the vertex coordinates in LDS,
// compute every distinct tet's quality once into LDS (its own tets are also
// written out, coalesced, and counted for the mesh), take every ball
// minimum from LDS, write one (min, sum) partial per tile.  A second kernel
// adds each mesh's tile partials in tile order (deterministic).
//
// The work-group count is the device's resident capacity (persistent
// kernel); each XCD walks a contiguous range of tiles so that neighbouring
// tiles, which share halo vertices, meet in the same L2.  The tile size TV
// and the ball padding Pad are chosen by timing every candidate on a sample
// of the batch's meshes (GmlxBatchTune).

#include "gml_internal.h"
#include "gml_device.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cstring>
#include <thread>

namespace gml {

static const int kBatBlock = 256;
static const int kTuneTV[] = {256, 512, 1024};
static const int kTunePad[] = {1, 2, 4};
static const int kNmbTune = 9;

// device arrays of one tile plan
struct BatchPlan {
   int TV = 0, Pad = 0;
   long long ntile = 0, nchunk = 0, sum_nv = 0, sum_nt = 0, nslab = 0;
   int max_nv = 0, max_nt = 0;
   unsigned int *vofs = nullptr;    // [ntile+1] tile vertex list ranges
   int *vlist = nullptr;            // tile vertex lists (global vertex ids)
   unsigned int *tofs = nullptr;    // [ntile+1] tile tet ranges
   v4u16 *tloc = nullptr;           // tile tets (own first): 4 tile-local vertex indices
   unsigned int *tbeg = nullptr;    // [ntile+1] first own tet of each tile (plan numbering)
   unsigned int *vbeg = nullptr;    // [ntile+1] own vertex ranges (global)
   unsigned int *cofs = nullptr;    // [ntile+1] first SELL-64 chunk of each tile
   unsigned int *loff = nullptr;    // [nchunk+1] slab offset of each chunk
   unsigned short *lball = nullptr; // ball slots: tile-local tet index (nt = sentinel)
   unsigned short *ldeg = nullptr;  // [V] padded ball length of each vertex
   unsigned int *mtile = nullptr;   // [nmesh+1] tile range of each mesh
   std::vector<int> perm;           // host: plan tet number -> batch tet number
   size_t bytes = 0;
};

struct Batch {
   int nmesh = 0;
   long long V = 0, T = 0;
   std::vector<long long> vstart, tstart;  // host: prefix sums of the sizes
   std::vector<int> htet;                  // host: local-index tets (plan builds)
   v4f *crd = nullptr;
   float *tq = nullptr, *vmin = nullptr;
   double *part = nullptr;                 // [2 * ntile]: tile min, tile sum
   float *mmin = nullptr;
   double *msum = nullptr;
   BatchPlan *plan = nullptr;
   int grid = 0;
   hipGraph_t graph = nullptr;
   hipGraphExec_t exec = nullptr;
   double tune[3 * kNmbTune] = {0};
   int ntune = 0;
   size_t bytes = 0;
};

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
template<int PAD>
__global__ void __launch_bounds__(kBatBlock) k_batch_quality(
   long long ntile, const unsigned int *__restrict__ vofs, const int *__restrict__ vlist,
   const unsigned int *__restrict__ tofs, const v4u16 *__restrict__ tloc, const unsigned int *__restrict__ tbeg,
   const unsigned int *__restrict__ vbeg, const unsigned int *__restrict__ cofs, const unsigned int *__restrict__ loff,
   const unsigned short *__restrict__ lball, const unsigned short *__restrict__ ldeg,
   const v4f *__restrict__ crd, float *__restrict__ tq, float *__restrict__ vmin, double *__restrict__ part, int max_nv)
{
   extern __shared__ __align__(16) unsigned char gml_lds[];
   v4f *X = reinterpret_cast<v4f *>(gml_lds);
   float *Q = reinterpret_cast<float *>(X + max_nv);
   __shared__ float rmin[kBatBlock / 64];
   __shared__ double rsum[kBatBlock / 64];

   // tiles of this XCD: a contiguous range, split among its work-groups
   const unsigned int x = blockIdx.x & 7, nper = gridDim.x >> 3, i0 = blockIdx.x >> 3;
   const long long t_lo = ntile * x / 8, t_hi = ntile * (x + 1) / 8;
   for(long long t = t_lo + i0; t < t_hi; t += nper)
   {
      const unsigned int v0 = vofs[t], nv = vofs[t + 1] - v0;
      const unsigned int t0 = tofs[t], nt = tofs[t + 1] - t0;
      // the tile's coordinates: four per thread and pass, the four index
      // loads, then the four gathers, in flight together
      for(unsigned int i = threadIdx.x; i < nv; i += 4 * kBatBlock)
      {
         int vi[4];
#pragma unroll
         for(int u = 0; u < 4; u++)
            vi[u] = i + u * kBatBlock < nv ? vlist[v0 + i + u * kBatBlock] : -1;
         v4f xv[4];
#pragma unroll
         for(int u = 0; u < 4; u++)
            xv[u] = vi[u] >= 0 ? crd[vi[u]] : (v4f){0.f, 0.f, 0.f, 0.f};
#pragma unroll
         for(int u = 0; u < 4; u++)
            if(vi[u] >= 0) X[i + u * kBatBlock] = xv[u];
      }
      if(threadIdx.x == 0) Q[nt] = FLT_MAX;
      __syncthreads();

      // tet qualities, four tets per thread and pass (their local rows
      // loaded together); each thread's own tets are summed in index order
      float mn = FLT_MAX;
      double sm = 0.;
      const unsigned int q0 = tbeg[t], nown = tbeg[t + 1] - q0;
      for(unsigned int j = threadIdx.x; j < nt; j += 4 * kBatBlock)
      {
         v4u16 a[4];
#pragma unroll
         for(int u = 0; u < 4; u++)
            a[u] = j + u * kBatBlock < nt ? tloc[t0 + j + u * kBatBlock] : (v4u16){0, 0, 0, 0};
#pragma unroll
         for(int u = 0; u < 4; u++)
         {
            const unsigned int jj = j + u * kBatBlock;
            if(jj < nt)
            {
               const float q = meanratio3(X[a[u].x], X[a[u].y], X[a[u].z], X[a[u].w]);
               Q[jj] = q;
               if(jj < nown)
               {
                  tq[q0 + jj] = q;
                  mn = fminf(mn, q);
                  sm += (double)q;
               }
            }
         }
      }
      __syncthreads();

      const unsigned int vb = vbeg[t], no = vbeg[t + 1] - vb, c0 = cofs[t];
      for(unsigned int i = threadIdx.x; i < no; i += kBatBlock)
      {
         const unsigned int v = vb + i;
         const int d = ldeg[v];
         const unsigned short *row = lball + (size_t)loff[c0 + (i >> 6)] * 64 + (i & 63);
         // the padded ball row in groups of 16 slots loaded together (a
         // slot past the row reads the sentinel, whose value is FLT_MAX)
         float m = FLT_MAX;
         for(int k = 0; k < d; k += 16)
         {
            unsigned short sl[16];
#pragma unroll
            for(int u = 0; u < 16; u++) sl[u] = k + u < d ? row[(k + u) * 64] : (unsigned short)nt;
#pragma unroll
            for(int u = 0; u < 16; u++) m = fminf(m, Q[sl[u]]);
         }
         vmin[v] = d > 0 ? m : 0.f;
      }

      // tile partials: min and sum over the owned tets
      for(int o = 32; o > 0; o >>= 1)
      {
         mn = fminf(mn, __shfl_xor(mn, o, 64));
         sm += __shfl_xor(sm, o, 64);
      }
      if((threadIdx.x & 63) == 0) { rmin[threadIdx.x >> 6] = mn; rsum[threadIdx.x >> 6] = sm; }
      __syncthreads();
      if(threadIdx.x == 0)
      {
         float a = rmin[0];
         double b = rsum[0];
         for(int w = 1; w < kBatBlock / 64; w++) { a = fminf(a, rmin[w]); b += rsum[w]; }
         part[2 * t] = (double)a;
         part[2 * t + 1] = b;
      }
      __syncthreads();
   }
}

// one wave per mesh: its tile partials in tile order
__global__ void __launch_bounds__(256) k_batch_finish(int nmesh, const unsigned int *__restrict__ mtile,
                                                      const double *__restrict__ part, float *__restrict__ mmin,
                                                      double *__restrict__ msum)
{
   const int m = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6), lane = threadIdx.x & 63;
   if(m >= nmesh) return;
   float a = FLT_MAX;
   double b = 0.;
   for(unsigned int t = mtile[m] + lane; t < mtile[m + 1]; t += 64)
   {
      a = fminf(a, (float)part[2 * t]);
      b += part[2 * t + 1];
   }
   for(int o = 32; o > 0; o >>= 1)
   {
      a = fminf(a, __shfl_xor(a, o, 64));
      b += __shfl_xor(b, o, 64);
   }
   if(lane == 0) { mmin[m] = a; msum[m] = b; }
}

// ---------------------------------------------------------------------------
// tile plan (host)
// ---------------------------------------------------------------------------
struct MeshTiles {
   std::vector<unsigned int> nvt, ntt, nownt, nck, nownt_t;  // per tile: vertices, tets, own vertices, chunks, own tets
   std::vector<int> vlist;                           // global vertex ids
   std::vector<unsigned short> tloc;                 // 4 per tile tet
   std::vector<int> perm;                            // own tets in plan order: mesh-local tet number
   std::vector<unsigned int> ckw;                    // chunk widths (slabs)
   std::vector<unsigned short> lball;                // chunk slabs, 64 lanes each
   std::vector<unsigned short> ldeg;                 // per own vertex (mesh order)
   bool ok = true;
};

// tiles of mesh m (tets with local vertex indices)
static void MeshPlan(const Batch *b, int m, int TV, int Pad, MeshTiles &o)
{
   const long long V = b->vstart[m + 1] - b->vstart[m], T = b->tstart[m + 1] - b->tstart[m];
   const int *tet = b->htet.data() + b->tstart[m] * 4;
   const long long vg = b->vstart[m];
   // vertex balls, ascending tet index
   std::vector<unsigned int> bo((size_t)V + 1, 0);
   for(long long e = 0; e < T * 4; e++) bo[tet[e] + 1]++;
   for(long long v = 0; v < V; v++) bo[v + 1] += bo[v];
   std::vector<int> ball(bo[V]);
   {
      std::vector<unsigned int> fill(bo.begin(), bo.end() - 1);
      for(long long e = 0; e < T; e++)
         for(int j = 0; j < 4; j++) ball[fill[tet[e * 4 + j]]++] = (int)e;
   }
   std::vector<int> mapv((size_t)V, -1), mapt((size_t)T, -1);
   std::vector<int> ts, vs, th;
   for(long long v0 = 0; v0 < V; v0 += TV)
   {
      const long long v1 = std::min(V, v0 + TV);
      ts.clear();
      for(long long v = v0; v < v1; v++)
         for(unsigned int k = bo[v]; k < bo[v + 1]; k++)
            if(mapt[ball[k]] < 0) { mapt[ball[k]] = 0; ts.push_back(ball[k]); }
      // own tets (node 0 in the tile) first, then the halo, each ascending
      std::sort(ts.begin(), ts.end());
      th.clear();
      size_t no = 0;
      for(int e : ts)
         if(tet[e * 4] >= v0 && tet[e * 4] < v1) ts[no++] = e;
         else th.push_back(e);
      std::copy(th.begin(), th.end(), ts.begin() + no);
      vs.clear();
      for(int e : ts)
         for(int j = 0; j < 4; j++)
            if(mapv[tet[e * 4 + j]] < 0) { mapv[tet[e * 4 + j]] = 0; vs.push_back(tet[e * 4 + j]); }
      std::sort(vs.begin(), vs.end());
      if(ts.size() > 65534 || vs.size() > 65535) { o.ok = false; return; }
      for(size_t i = 0; i < vs.size(); i++) { mapv[vs[i]] = (int)i; o.vlist.push_back((int)(vg + vs[i])); }
      for(size_t i = 0; i < ts.size(); i++)
      {
         const int e = ts[i];
         mapt[e] = (int)i;
         for(int j = 0; j < 4; j++) o.tloc.push_back((unsigned short)mapv[tet[e * 4 + j]]);
         if(i < no) o.perm.push_back(e);
      }
      o.nownt_t.push_back((unsigned int)no);
      // own vertices' balls in SELL-64 chunks, rows padded to Pad with the sentinel
      const unsigned short sent = (unsigned short)ts.size();
      unsigned int nck = 0;
      for(long long c0 = v0; c0 < v1; c0 += 64, nck++)
      {
         const long long c1 = std::min(v1, c0 + 64);
         unsigned int w = 0;
         for(long long v = c0; v < c1; v++)
         {
            unsigned int d = bo[v + 1] - bo[v];
            d = (d + Pad - 1) / Pad * Pad;
            o.ldeg.push_back((unsigned short)d);
            w = std::max(w, d);
         }
         o.ckw.push_back(w);
         const size_t base = o.lball.size();
         o.lball.resize(base + (size_t)w * 64, sent);
         for(long long v = c0; v < c1; v++)
            for(unsigned int k = bo[v]; k < bo[v + 1]; k++)
               o.lball[base + (size_t)(k - bo[v]) * 64 + (size_t)(v - c0)] = (unsigned short)mapt[ball[k]];
      }
      o.nvt.push_back((unsigned int)vs.size());
      o.ntt.push_back((unsigned int)ts.size());
      o.nownt.push_back((unsigned int)(v1 - v0));
      o.nck.push_back(nck);
      for(int e : ts) mapt[e] = -1;
      for(int v : vs) mapv[v] = -1;
   }
}

static void FreePlan(GmlCtx *g, BatchPlan *p)
{
   if(!p) return;
   DevFree(g, p->vofs); DevFree(g, p->vlist); DevFree(g, p->tofs); DevFree(g, p->tloc); DevFree(g, p->tbeg);
   DevFree(g, p->vbeg); DevFree(g, p->cofs); DevFree(g, p->loff); DevFree(g, p->lball); DevFree(g, p->ldeg);
   DevFree(g, p->mtile);
   g->MemSiz -= p->bytes;
   delete p;
}

// plan over the meshes listed in ms (ascending), their vertices renumbered
// contiguously in that order when ms is a sample of the batch
static BatchPlan *BuildPlan(GmlCtx *g, const Batch *b, const std::vector<int> &ms, int TV, int Pad)
{
   const int nm = (int)ms.size();
   std::vector<MeshTiles> mt(nm);
   std::atomic<int> next(0);
   unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
   std::vector<std::thread> th;
   for(unsigned w = 0; w < nth; w++)
      th.emplace_back([&]() {
         for(int i; (i = next++) < nm;) MeshPlan(b, ms[i], TV, Pad, mt[i]);
      });
   for(auto &x : th) x.join();
   for(auto &o : mt) if(!o.ok) return nullptr;

   BatchPlan *p = new BatchPlan();
   p->TV = TV; p->Pad = Pad;
   std::vector<unsigned int> vofs(1, 0), tofs(1, 0), vbeg(1, 0), cofs(1, 0), loff(1, 0), mtile(1, 0), tbeg(1, 0);
   size_t nl = 0, ntl = 0, nb = 0, nd = 0;
   for(int i = 0; i < nm; i++)
   {
      auto &o = mt[i];
      for(size_t k = 0; k < o.nvt.size(); k++)
      {
         vofs.push_back(vofs.back() + o.nvt[k]);
         tofs.push_back(tofs.back() + o.ntt[k]);
         vbeg.push_back(vbeg.back() + o.nownt[k]);
         cofs.push_back(cofs.back() + o.nck[k]);
         tbeg.push_back(tbeg.back() + o.nownt_t[k]);
         p->max_nv = std::max(p->max_nv, (int)o.nvt[k]);
         p->max_nt = std::max(p->max_nt, (int)o.ntt[k]);
      }
      for(unsigned int w : o.ckw) loff.push_back(loff.back() + w);
      mtile.push_back((unsigned int)(vofs.size() - 1));
      nl += o.vlist.size(); ntl += o.tloc.size() / 4; nb += o.lball.size(); nd += o.ldeg.size();
   }
   p->ntile = (long long)vofs.size() - 1;
   p->nchunk = (long long)loff.size() - 1;
   p->nslab = loff.back();
   p->sum_nv = vofs.back();
   p->sum_nt = tofs.back();
   // concatenate (vertex ids of a sample are renumbered to its own prefix)
   std::vector<int> vlist(nl + 1);
   p->perm.resize(tbeg.back());
   std::vector<unsigned short> tloc(ntl * 4 + 4), lball(nb + 64), ldeg(nd + 1);
   {
      size_t a = 0, c = 0, e = 0, f = 0, q = 0;
      long long vshift = 0, tshift = 0;
      for(int i = 0; i < nm; i++)
      {
         auto &o = mt[i];
         const long long dv = b->vstart[ms[i]] - vshift;
         for(int v : o.vlist) vlist[a++] = (int)(v - dv);
         // own tets in plan order -> their number in the (sample's) batch order
         for(int x : o.perm) p->perm[q++] = (int)(tshift + x);
         std::copy(o.tloc.begin(), o.tloc.end(), tloc.begin() + c * 4);
         c += o.tloc.size() / 4;
         std::copy(o.lball.begin(), o.lball.end(), lball.begin() + e);
         e += o.lball.size();
         std::copy(o.ldeg.begin(), o.ldeg.end(), ldeg.begin() + f);
         f += o.ldeg.size();
         vshift += b->vstart[ms[i] + 1] - b->vstart[ms[i]];
         tshift += b->tstart[ms[i] + 1] - b->tstart[ms[i]];
      }
   }
   bool ok = true;
   auto up = [&](void **d, const void *h, size_t by) {
      *d = DevAlloc(g, by);
      if(!*d || !DevCopy(g, *d, h, by, 1)) ok = false;
      p->bytes += by;
   };
   up((void **)&p->vofs, vofs.data(), vofs.size() * 4);
   up((void **)&p->tofs, tofs.data(), tofs.size() * 4);
   up((void **)&p->vbeg, vbeg.data(), vbeg.size() * 4);
   up((void **)&p->cofs, cofs.data(), cofs.size() * 4);
   up((void **)&p->loff, loff.data(), loff.size() * 4);
   up((void **)&p->mtile, mtile.data(), mtile.size() * 4);
   up((void **)&p->vlist, vlist.data(), vlist.size() * 4);
   up((void **)&p->tbeg, tbeg.data(), tbeg.size() * 4);
   up((void **)&p->tloc, tloc.data(), tloc.size() * 2);
   up((void **)&p->lball, lball.data(), lball.size() * 2);
   up((void **)&p->ldeg, ldeg.data(), ldeg.size() * 2);
   g->MemSiz += p->bytes;
   if(!ok) { FreePlan(g, p); return nullptr; }
   return p;
}

static size_t PlanLds(const BatchPlan *p) { return (size_t)p->max_nv * 16 + ((size_t)p->max_nt + 1) * 4; }

// resident work-groups of the persistent kernel: a multiple of 8 (one
// share per XCD), at most 8 per XCD more than there are tiles
static int PersistentGrid(GmlCtx *g, const BatchPlan *p, const void *kern)
{
   int ncu = 0, per = 0;
   if(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, g->dev) != hipSuccess || ncu < 1) ncu = 256;
   if(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, kBatBlock, PlanLds(p)) != hipSuccess || per < 1) per = 1;
   long long grid = (long long)ncu * per;
   grid = std::min<long long>(grid, (p->ntile + 7) / 8 * 8);
   grid = std::max<long long>(8, grid / 8 * 8);
   return (int)grid;
}

static const void *QualityKernel(int Pad)
{
   if(Pad == 1) return (const void *)k_batch_quality<1>;
   if(Pad == 2) return (const void *)k_batch_quality<2>;
   return (const void *)k_batch_quality<4>;
}

// allow the plan's LDS size (before any launch or stream capture)
static bool PrepareQuality(const BatchPlan *p)
{
   const size_t lds = PlanLds(p);
   if(lds > 160 * 1024) return false;
   if(lds > 64 * 1024
   && hipFuncSetAttribute(QualityKernel(p->Pad), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
      return false;
   return true;
}

static bool LaunchQuality(GmlCtx *g, const BatchPlan *p, const v4f *crd, float *tq, float *vmin, double *part, int grid)
{
   const size_t lds = PlanLds(p);
#define GML_Q(P) hipLaunchKernelGGL(k_batch_quality<P>, dim3(grid), dim3(kBatBlock), lds, g->stream, p->ntile, p->vofs, \
      p->vlist, p->tofs, p->tloc, p->tbeg, p->vbeg, p->cofs, p->loff, p->lball, p->ldeg, crd, tq, vmin, part, p->max_nv)
   if(p->Pad == 1) GML_Q(1);
   else if(p->Pad == 2) GML_Q(2);
   else GML_Q(4);
#undef GML_Q
   return hipGetLastError() == hipSuccess;
}

static void DropGraph(Batch *b)
{
   if(b->exec) (void)hipGraphExecDestroy(b->exec);
   if(b->graph) (void)hipGraphDestroy(b->graph);
   b->exec = nullptr;
   b->graph = nullptr;
}

static void FreeBatch(GmlCtx *g, Batch *b)
{
   if(!b) return;
   DropGraph(b);
   FreePlan(g, b->plan);
   DevFree(g, b->crd); DevFree(g, b->tq); DevFree(g, b->vmin); DevFree(g, b->part);
   DevFree(g, b->mmin); DevFree(g, b->msum);
   g->MemSiz -= b->bytes;
   delete b;
}

static Batch *Bat(GmlCtx *g, int i) { return (i >= 1 && i <= GmlxMaxBatch) ? g->bat[i] : nullptr; }

void FreeBatches(GmlCtx *g)
{
   for(int i = 1; i <= GmlxMaxBatch; i++)
   {
      FreeBatch(g, g->bat[i]);
      g->bat[i] = nullptr;
   }
}

// install plan p as the batch's plan: partial buffer, grid, graph
static int UsePlan(GmlCtx *g, Batch *b, BatchPlan *p)
{
   DropGraph(b);
   FreePlan(g, b->plan);
   b->plan = p;
   DevFree(g, b->part);
   b->part = (double *)DevAlloc(g, (size_t)p->ntile * 16 + 16);
   if(!b->part) return -5;
   if(!PrepareQuality(p)) return -5;
   b->grid = PersistentGrid(g, p, QualityKernel(p->Pad));
   // capture the two kernels once; every run replays the graph
   if(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return -6;
   bool ok = LaunchQuality(g, p, b->crd, b->tq, b->vmin, b->part, b->grid);
   if(ok)
   {
      hipLaunchKernelGGL(k_batch_finish, dim3((b->nmesh + 3) / 4), dim3(256), 0, g->stream, b->nmesh, p->mtile,
                         b->part, b->mmin, b->msum);
      ok = hipGetLastError() == hipSuccess;
   }
   hipGraph_t gr = nullptr;
   if(hipStreamEndCapture(g->stream, &gr) != hipSuccess || !ok)
   {
      if(gr) (void)hipGraphDestroy(gr);
      return -6;
   }
   b->graph = gr;
   if(hipGraphInstantiate(&b->exec, gr, nullptr, nullptr, 0) != hipSuccess)
   {
      DropGraph(b);
      return -6;
   }
   DbgPrint(g, "Batch plan: %d meshes, %lld tiles of %d vertices, ball padding %d, %lld tile tets (%.2f per tet), "
            "%lld tile vertices (%.2f per vertex), LDS %zu B, grid %d, %zu MB\n", b->nmesh, p->ntile, p->TV, p->Pad,
            p->sum_nt, (double)p->sum_nt / (double)b->T, p->sum_nv, (double)p->sum_nv / (double)b->V, PlanLds(p),
            b->grid, p->bytes >> 20);
   return 1;
}

static bool ValidShape(int TV, int Pad)
{
   return TV >= 64 && TV <= 1024 && !(TV & 63) && (Pad == 1 || Pad == 2 || Pad == 4);
}

} // namespace gml

using namespace gml;

extern "C" {

int GmlxNewBatch(size_t GmlIdx, int NmbMsh, const int *NmbVer, const int *NmbTet, const float *crd, const int *tet)
{
   GML_CTX(g, GmlIdx);
   if(g->host_only || NmbMsh < 1 || !NmbVer || !NmbTet || !crd || !tet) return 0;
   int idx = 0;
   for(int i = 1; i <= GmlxMaxBatch; i++) if(!g->bat[i]) { idx = i; break; }
   if(!idx) return 0;
   Batch *b = new Batch();
   b->nmesh = NmbMsh;
   b->vstart.assign(NmbMsh + 1, 0);
   b->tstart.assign(NmbMsh + 1, 0);
   for(int m = 0; m < NmbMsh; m++)
   {
      if(NmbVer[m] < 1 || NmbTet[m] < 0) { delete b; return 0; }
      b->vstart[m + 1] = b->vstart[m] + NmbVer[m];
      b->tstart[m + 1] = b->tstart[m] + NmbTet[m];
   }
   b->V = b->vstart[NmbMsh];
   b->T = b->tstart[NmbMsh];
   if(b->V >= (1LL << 31) - 1 || b->T >= (1LL << 31) - 1) { delete b; return 0; }
   b->htet.assign(tet, tet + b->T * 4);
   for(int m = 0; m < NmbMsh; m++)
      for(long long e = b->tstart[m] * 4; e < b->tstart[m + 1] * 4; e++)
         if(b->htet[e] < 0 || b->htet[e] >= NmbVer[m]) { delete b; return 0; }
   std::vector<float> c(crd, crd + b->V * 4);
   for(long long v = 0; v < b->V; v++) c[v * 4 + 3] = 0.f;
   b->crd = (v4f *)DevAlloc(g, (size_t)b->V * 16);
   b->tq = (float *)DevAlloc(g, (size_t)b->T * 4 + 4);
   b->vmin = (float *)DevAlloc(g, (size_t)b->V * 4);
   b->mmin = (float *)DevAlloc(g, (size_t)NmbMsh * 4);
   b->msum = (double *)DevAlloc(g, (size_t)NmbMsh * 8);
   b->bytes = (size_t)b->V * 20 + (size_t)b->T * 4 + 4 + (size_t)NmbMsh * 12;
   g->MemSiz += b->bytes;
   if(!b->crd || !b->tq || !b->vmin || !b->mmin || !b->msum || !DevCopy(g, b->crd, c.data(), (size_t)b->V * 16, 1))
   {
      FreeBatch(g, b);
      return 0;
   }
   g->MovSiz += (size_t)b->V * 16;
   g->bat[idx] = b;
   return idx;
}

int GmlxBatchSetShape(size_t GmlIdx, int BatIdx, int TV, int Pad)
{
   GML_CTX(g, GmlIdx);
   Batch *b = Bat(g, BatIdx);
   if(!b) return -1;
   if(!ValidShape(TV, Pad)) return -2;
   std::vector<int> all(b->nmesh);
   for(int m = 0; m < b->nmesh; m++) all[m] = m;
   BatchPlan *p = BuildPlan(g, b, all, TV, Pad);
   if(!p) return -5;
   return UsePlan(g, b, p);
}

int GmlxBatchTune(size_t GmlIdx, int BatIdx, int *TV, int *Pad)
{
   GML_CTX(g, GmlIdx);
   Batch *b = Bat(g, BatIdx);
   if(!b) return -1;
   // sample: every k-th mesh, k chosen for about 1/16 of the meshes
   std::vector<int> ms;
   const int step = std::max(1, b->nmesh / 16);
   for(int m = 0; m < b->nmesh; m += step) ms.push_back(m);
   long long sv = 0, st = 0;
   for(int m : ms) { sv += b->vstart[m + 1] - b->vstart[m]; st += b->tstart[m + 1] - b->tstart[m]; }
   float *tq = (float *)DevAlloc(g, (size_t)st * 4 + 4), *vm = (float *)DevAlloc(g, (size_t)sv * 4);
   v4f *crd = (v4f *)DevAlloc(g, (size_t)sv * 16);
   bool ok = tq && vm && crd;
   {
      long long o = 0;
      for(int m : ms)
      {
         const long long n = b->vstart[m + 1] - b->vstart[m];
         if(ok && !DevCopy(g, crd + o, b->crd + b->vstart[m], (size_t)n * 16, 3)) ok = false;
         o += n;
      }
   }
   double best = 1e30;
   int bt = 512, bp = 1;
   b->ntune = 0;
   hipEvent_t e0 = nullptr, e1 = nullptr;
   if(ok && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) ok = false;
   for(int i = 0; ok && i < 3; i++)
      for(int j = 0; ok && j < 3; j++)
      {
         BatchPlan *p = BuildPlan(g, b, ms, kTuneTV[i], kTunePad[j]);
         if(!p) continue;
         double *part = (double *)DevAlloc(g, (size_t)p->ntile * 16 + 16);
         const bool prep = PrepareQuality(p);
         const int grid = PersistentGrid(g, p, QualityKernel(p->Pad));
         double us = -1.;
         if(part && prep && LaunchQuality(g, p, crd, tq, vm, part, grid))
         {
            float ms_ = 0.f;
            (void)hipEventRecord(e0, g->stream);
            for(int r = 0; r < 3; r++) LaunchQuality(g, p, crd, tq, vm, part, grid);
            (void)hipEventRecord(e1, g->stream);
            if(hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms_, e0, e1) == hipSuccess)
               us = ms_ * 1e3 / 3.;
         }
         b->tune[3 * b->ntune] = p->TV;
         b->tune[3 * b->ntune + 1] = p->Pad;
         b->tune[3 * b->ntune + 2] = us;
         b->ntune++;
         if(us > 0 && us < best) { best = us; bt = p->TV; bp = p->Pad; }
         DevFree(g, part);
         FreePlan(g, p);
      }
   if(e0) (void)hipEventDestroy(e0);
   if(e1) (void)hipEventDestroy(e1);
   DevFree(g, tq); DevFree(g, vm); DevFree(g, crd);
   if(!ok) return -5;
   DbgPrint(g, "Batch tuner: %d candidates on %zu of %d meshes, best tile %d, padding %d (%.1f us)\n",
            b->ntune, ms.size(), b->nmesh, bt, bp, best);
   if(TV) *TV = bt;
   if(Pad) *Pad = bp;
   return GmlxBatchSetShape(GmlIdx, BatIdx, bt, bp);
}

int GmlxBatchQuality(size_t GmlIdx, int BatIdx, float *MshMin, double *MshMean)
{
   GML_CTX(g, GmlIdx);
   Batch *b = Bat(g, BatIdx);
   if(!b) return -1;
   if(!b->exec)
   {
      int rc = GmlxBatchTune(GmlIdx, BatIdx, nullptr, nullptr);
      if(rc != 1) return rc;
   }
   int rc = TimedLaunch(g, GmlxBuiltinBatch, [&]() { (void)hipGraphLaunch(b->exec, g->stream); });
   if(rc != 1) return rc;
   std::vector<double> sum(b->nmesh);
   if((MshMin && !DevCopy(g, MshMin, b->mmin, (size_t)b->nmesh * 4, 2))
   || !DevCopy(g, sum.data(), b->msum, (size_t)b->nmesh * 8, 2))
      return -6;
   if(MshMean)
      for(int m = 0; m < b->nmesh; m++)
      {
         const long long n = b->tstart[m + 1] - b->tstart[m];
         MshMean[m] = n ? sum[m] / (double)n : 0.;
      }
   return 1;
}

int GmlxGetBatchField(size_t GmlIdx, int BatIdx, float *TetQal, float *VerMin)
{
   GML_CTX(g, GmlIdx);
   Batch *b = Bat(g, BatIdx);
   if(!b) return -1;
   if(TetQal)
   {
      if(!b->plan) return -1;
      std::vector<float> q((size_t)b->T);
      if(!DevCopy(g, q.data(), b->tq, (size_t)b->T * 4, 2)) return -6;
      const std::vector<int> &pm = b->plan->perm;
      for(long long i = 0; i < b->T; i++) TetQal[pm[i]] = q[i];
   }
   if(VerMin && !DevCopy(g, VerMin, b->vmin, (size_t)b->V * 4, 2)) return -6;
   return 1;
}

int GmlxGetBatchInfo(size_t GmlIdx, int BatIdx, long long *info)
{
   GML_CTX(g, GmlIdx);
   Batch *b = Bat(g, BatIdx);
   if(!b || !info) return 0;
   const BatchPlan *p = b->plan;
   long long v[12] = {b->nmesh, b->V, b->T, p ? p->ntile : 0, p ? p->TV : 0, p ? p->Pad : 0,
                      p ? (long long)p->bytes : 0, p ? p->sum_nt : 0, p ? p->sum_nv : 0, b->grid,
                      b->exec ? 1 : 0, p ? p->nslab : 0};
   memcpy(info, v, sizeof(v));
   return 1;
}

int GmlxGetBatchTuning(size_t GmlIdx, int BatIdx, double *tab)
{
   GML_CTX(g, GmlIdx);
   Batch *b = Bat(g, BatIdx);
   if(!b || !tab) return 0;
   memcpy(tab, b->tune, sizeof(double) * 3 * b->ntune);
   return b->ntune;
}

int GmlxFreeBatch(size_t GmlIdx, int BatIdx)
{
   GML_CTX(g, GmlIdx);
   Batch *b = Bat(g, BatIdx);
   if(!b) return 0;
   FreeBatch(g, b);
   g->bat[BatIdx] = nullptr;
   return 1;
}

} // extern "C"
