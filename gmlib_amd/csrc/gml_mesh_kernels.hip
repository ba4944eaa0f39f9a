// Hand-written HIP kernels for the mesh workloads of this engine: tet mean-
// ratio quality, per-vertex ball minimum, and Jacobi barycentre smoothing
// (tet scatter + vertex-ball gather).  They compute exactly what the
// generated kernels built from gmlib_amd/kernels/*.cl compute, expression by
// expression, so their results are bit-identical to the drop-in path (the
// quality kernel drops the terms of the coordinates' fourth component, which
// the C API always stores as 0, see meanratio3); what differs is how they
// move data:
//   - element rows are streamed with non-temporal loads (read once),
//   - the vertex ball is read from its SELL-64 slabs with one 256-byte load
//     per slot and wave, and the barycentre gathers of a row are issued in
//     batches of 8 before any is consumed,
//   - blocks are remapped so that each XCD (own 4 MB L2) walks a contiguous
//     range of the Hilbert-ordered mesh.
// Reference patterns: examples/MeshQuality (quality), examples/BallQuality
// (ball min), examples/NodeSmoothing (scatter + gather).

#include "gml_internal.h"
#include "gml_device.h"

namespace gml {

// per-handle timing records of the built-in kernels (index = GmlxBuiltin*)
static std::mutex g_stat_mtx;
static std::vector<std::pair<GmlCtx *, KernStat *>> g_stats;

KernStat *Stat(GmlCtx *g, int which)
{
   std::lock_guard<std::mutex> lk(g_stat_mtx);
   for(auto &p : g_stats)
      if(p.first == g) return &p.second[which];
   KernStat *s = new KernStat[kNmbBuiltin];
   g_stats.push_back({g, s});
   return &s[which];
}

void FreeBuiltinStats(GmlCtx *g)
{
   std::lock_guard<std::mutex> lk(g_stat_mtx);
   for(size_t i = 0; i < g_stats.size(); i++)
      if(g_stats[i].first == g)
      {
         for(int w = 0; w < kNmbBuiltin; w++)
            for(auto &e : g_stats[i].second[w].ev) FreeRec(e);
         delete[] g_stats[i].second;
         g_stats.erase(g_stats.begin() + i);
         return;
      }
}

// TPT tets per thread (a work-group covers TPT x its size, each part
// coalesced): 4 TPT coordinate gathers in flight per thread; a missing tet
// of the last group repeats the thread's first one and is not stored
template <int TPT>
__global__ void __launch_bounds__(1024) k_tet_quality(long long nt, const v4i *__restrict__ tet,
                                                       const v4f *__restrict__ crd, float *__restrict__ q, int remap)
{
   const long long t0 = (long long)xcd_block(remap) * TPT * blockDim.x + threadIdx.x;
   if(t0 >= nt) return;
   v4i e[TPT];
   e[0] = ld_nt(&tet[t0]);
#pragma unroll
   for(int u = 1; u < TPT; u++)
      e[u] = t0 + u * blockDim.x < nt ? ld_nt(&tet[t0 + u * blockDim.x]) : e[0];
   v4f c[TPT][4];
#pragma unroll
   for(int u = 0; u < TPT; u++)
   {
      c[u][0] = crd[e[u].x]; c[u][1] = crd[e[u].y]; c[u][2] = crd[e[u].z]; c[u][3] = crd[e[u].w];
   }
#pragma unroll
   for(int u = 0; u < TPT; u++)
      if(t0 + u * blockDim.x < nt)
         __builtin_nontemporal_store(meanratio3(c[u][0], c[u][1], c[u][2], c[u][3]), &q[t0 + u * blockDim.x]);
}

// Tile-staged form: a work-group takes one tile of the element tile plan
// (R * blockDim.x consecutive tets, R per thread), loads the
// tile's distinct vertices into LDS (ascending ids: coalesced, two loads in
// flight per thread), and each tet reads its corners there through its four
// one-byte tile-local indices: 4 bytes per tet instead of its 16-byte row
// and four scattered 16-byte gathers.
template <int R>
__global__ void __launch_bounds__(256) k_tet_quality_tile(long long nt, const unsigned int *__restrict__ tofs,
                                                           const int *__restrict__ tlist, const unsigned int *__restrict__ lid,
                                                           const v4f *__restrict__ crd, float *__restrict__ q, int remap)
{
   __shared__ v4f s[256];
   const unsigned int t = xcd_block(remap);
   const long long base = (long long)t * R * blockDim.x;
   unsigned int li[R];
#pragma unroll
   for(int r = 0; r < R; r++)
      li[r] = __builtin_nontemporal_load(&lid[min(base + r * (long long)blockDim.x + threadIdx.x, nt - 1)]);
   const unsigned int t0 = tofs[t], n = tofs[t + 1] - t0;
   for(unsigned int i = threadIdx.x; i < n; i += 2 * blockDim.x)
   {
      const unsigned int j1 = min(i + blockDim.x, n - 1);
      const int a = tlist[t0 + i], b = tlist[t0 + j1];
      const v4f va = crd[a], vb = crd[b];
      s[i] = va;
      s[j1] = vb;
   }
   __syncthreads();
#pragma unroll
   for(int r = 0; r < R; r++)
   {
      const long long i = base + r * (long long)blockDim.x + threadIdx.x;
      const unsigned int w = li[r];
      if(i < nt)
         __builtin_nontemporal_store(meanratio3(s[w & 255u], s[(w >> 8) & 255u], s[(w >> 16) & 255u], s[w >> 24]), &q[i]);
   }
}

__global__ void __launch_bounds__(1024) k_smooth_scatter(long long nt, const v4i *__restrict__ tet,
                                                          const v4f *__restrict__ crd, v4f *__restrict__ bary, int remap)
{
   long long t = (long long)xcd_block(remap) * blockDim.x + threadIdx.x;
   if(t >= nt) return;
   v4i v = ld_nt(&tet[t]);
   v4f a = crd[v.x], b = crd[v.y], c = crd[v.z], d = crd[v.w];
   bary[t] = (((a + b) + c) + d) * 0.25f;
}

// Ball gather: slot k of vertex v is ent[(off[v >> 6] + k) * 64 + (v & 63)].
__global__ void __launch_bounds__(1024) k_smooth_gather(long long nv, const int *__restrict__ deg,
                                                         const unsigned int *__restrict__ off, const int *__restrict__ ent,
                                                         const v4f *__restrict__ bary, const int *__restrict__ ref,
                                                         v4f *__restrict__ crd, float omega, int remap)
{
   long long v = (long long)xcd_block(remap) * blockDim.x + threadIdx.x;
   if(v >= nv) return;
   const int d = deg[v];
   const int *row = ent + (size_t)off[v >> 6] * 64 + (v & 63);
   v4f s = {0.f, 0.f, 0.f, 0.f};
   for(int k = 0; k < d; k += 8)
   {
      int e[8];
#pragma unroll
      for(int i = 0; i < 8; i++) e[i] = k + i < d ? __builtin_nontemporal_load(&row[(k + i) * 64]) >> 4 : -1;
      v4f b[8];
#pragma unroll
      for(int i = 0; i < 8; i++) b[i] = e[i] >= 0 ? bary[e[i]] : (v4f){0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for(int i = 0; i < 8; i++) if(e[i] >= 0) s += b[i];
   }
   const int r = ref[v];
   if(!r && d > 0)
   {
      v4f p = crd[v];
      s *= 1.f / (float)d;
      crd[v] = p + (s - p) * omega;
   }
}

__global__ void __launch_bounds__(1024) k_ball_min(long long nv, const int *__restrict__ deg,
                                                    const unsigned int *__restrict__ off, const int *__restrict__ ent,
                                                    const float *__restrict__ val, float *__restrict__ out, int remap)
{
   long long v = (long long)xcd_block(remap) * blockDim.x + threadIdx.x;
   if(v >= nv) return;
   const int d = deg[v];
   const int *row = ent + (size_t)off[v >> 6] * 64 + (v & 63);
   // slots 8 at a time: the 8 entry loads, then the 8 value gathers.  NaN
   // is fminf's neutral element (fminf(NaN, x) = x), so starting from it and
   // padding with it gives the first value's then the running minimum, as
   // a loop from slot 0
   const float nanv = __builtin_nanf("");
   float m = nanv;
   for(int k = 0; k < d; k += 8)
   {
      int e[8];
#pragma unroll
      for(int i = 0; i < 8; i++) e[i] = k + i < d ? __builtin_nontemporal_load(&row[(k + i) * 64]) >> 4 : -1;
      float x[8];
#pragma unroll
      for(int i = 0; i < 8; i++) x[i] = e[i] >= 0 ? val[e[i]] : nanv;
#pragma unroll
      for(int i = 0; i < 8; i++) m = fminf(m, x[i]);
   }
   out[v] = d > 0 ? m : 0.f;
}

// Fused Jacobi sweep over one tile of TV = blockDim.x vertices (see the
// tile plan, gml_plan.cpp): stage the tile's vertex coordinates, form its
// distinct barycentres, gather every ball from LDS, write the new
// coordinates to the other buffer.
__global__ void __launch_bounds__(256) k_sweep_fused(long long nv, const unsigned int *__restrict__ vofs,
                                                      const int *__restrict__ vlist, const unsigned int *__restrict__ tofs,
                                                      const v4u16 *__restrict__ tloc, const int *__restrict__ deg,
                                                      const unsigned int *__restrict__ off, const unsigned short *__restrict__ lidx,
                                                      const int *__restrict__ ref, const v4f *__restrict__ xo,
                                                      v4f *__restrict__ xn, float omega, int remap)
{
   extern __shared__ v4f lds[];
   const unsigned int tile = xcd_block(remap);
   const int tid = threadIdx.x, TV = blockDim.x;
   const unsigned int v0 = vofs[tile], nvt = vofs[tile + 1] - v0;
   const unsigned int t0 = tofs[tile], ntt = tofs[tile + 1] - t0;
   v4f *X = lds, *Bc = lds + nvt;
   // 1. tile coordinates: index loads first, then the gathers, 4 in flight
   for(unsigned int i = tid; i < nvt; i += 4 * TV)
   {
      int vi[4];
#pragma unroll
      for(int u = 0; u < 4; u++)
         vi[u] = i + u * TV < nvt ? __builtin_nontemporal_load(&vlist[v0 + i + u * TV]) : -1;
      v4f x[4];
#pragma unroll
      for(int u = 0; u < 4; u++)
         x[u] = vi[u] >= 0 ? xo[vi[u]] : (v4f){0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for(int u = 0; u < 4; u++)
         if(vi[u] >= 0) X[i + u * TV] = x[u];
   }
   __syncthreads();
   // 2. distinct barycentres, in the scatter kernel's operation order
   for(unsigned int j = tid; j < ntt; j += 4 * TV)
   {
      v4u16 l[4];
#pragma unroll
      for(int u = 0; u < 4; u++)
         l[u] = j + u * TV < ntt ? tloc[t0 + j + u * TV] : (v4u16){0, 0, 0, 0};
#pragma unroll
      for(int u = 0; u < 4; u++)
         if(j + u * TV < ntt)
            Bc[j + u * TV] = (((X[l[u].x] + X[l[u].y]) + X[l[u].z]) + X[l[u].w]) * 0.25f;
   }
   __syncthreads();
   // 3. ball sums from LDS in ball order, slot indices loaded 8 at a time
   const long long v = (long long)tile * TV + tid;
   if(v >= nv) return;
   const int d = deg[v];
   const unsigned short *row = lidx + (size_t)off[v >> 6] * 64 + (v & 63);
   v4f s = {0.f, 0.f, 0.f, 0.f};
   for(int k = 0; k < d; k += 8)
   {
      int li[8];
#pragma unroll
      for(int u = 0; u < 8; u++)
         li[u] = k + u < d ? (int)__builtin_nontemporal_load(&row[(k + u) * 64]) : -1;
#pragma unroll
      for(int u = 0; u < 8; u++)
         if(li[u] >= 0) s += Bc[li[u]];
   }
   v4f p = xo[v];
   if(!ref[v] && d > 0)
   {
      s *= 1.f / (float)d;
      p = p + (s - p) * omega;
   }
   xn[v] = p;
}

// Operator-form Jacobi sweep over one tile: stage the tile's 1-ring
// coordinates in LDS (its own vertices at local 0..TV-1 straight from the
// coordinate array, the listed halo from TV on), then
// x_v += omega (sum_u c_vu x_u / (4 deg v) - x_v) with the 1-ring read as
// (tile-local index << 4 | c_vu) slots.  Every lane reads its SELL-64
// slab to the slab's width (zero slots past its ring are skipped), and
// deg v = (sum of its counts) / 3: each ball tet of v contributes its three
// other vertex positions (a repeat of v among them is a self slot).  The
// vertex's own data (ref, coordinates, first 16 ring slots) are loaded
// before the halo staging so that both latencies overlap.  One vertex per
// thread, TV = work-group size (64..1024).
__global__ void __launch_bounds__(1024) k_sweep_op(long long nv, const unsigned int *__restrict__ vofs,
                                                    const int *__restrict__ vlist,
                                                    const unsigned int *__restrict__ noff,
                                                    const unsigned short *__restrict__ nbr, const int *__restrict__ ref,
                                                    const v4f *__restrict__ xo, v4f *__restrict__ xn, float omega,
                                                    int remap)
{
   extern __shared__ v4f lds[];
   const unsigned int tile = xcd_block(remap);
   const int tid = threadIdx.x, TV = blockDim.x;
   const long long v = (long long)tile * TV + tid;
   const bool act = v < nv;
   int m = 0;
   size_t o = 0;
   if(act)
   {
      o = noff[v >> 6];
      m = (int)(noff[(v >> 6) + 1] - o);
   }
   const unsigned short *row = nbr + (act ? o * 64 + (v & 63) : 0);
   unsigned int e0[16];
#pragma unroll
   for(int u = 0; u < 16; u++)
      e0[u] = u < m ? __builtin_nontemporal_load(&row[u * 64]) : 0u;
   v4f p = act ? xo[v] : (v4f){0.f, 0.f, 0.f, 0.f};
   const int rf = act ? ref[v] : 1;

   const unsigned int h0 = vofs[tile], nh = vofs[tile + 1] - h0;
   v4f *X = lds;
   X[tid] = p;
   v4f *H = lds + TV;
   for(unsigned int i = tid; i < nh; i += 4 * TV)
   {
      int vi[4];
#pragma unroll
      for(int u = 0; u < 4; u++)
         vi[u] = i + u * TV < nh ? __builtin_nontemporal_load(&vlist[h0 + i + u * TV]) : -1;
      v4f x[4];
#pragma unroll
      for(int u = 0; u < 4; u++)
         x[u] = vi[u] >= 0 ? xo[vi[u]] : (v4f){0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for(int u = 0; u < 4; u++)
         if(vi[u] >= 0) H[i + u * TV] = x[u];
   }
   __syncthreads();
   if(!act) return;
   v4f s = {0.f, 0.f, 0.f, 0.f};
   int cs = 0;
#pragma unroll
   for(int u = 0; u < 16; u++)
      if(u < m && e0[u])
      {
         s += X[e0[u] >> 4] * (float)(e0[u] & 15u);
         cs += (int)(e0[u] & 15u);
      }
   for(int k = 16; k < m; k += 8)
   {
      unsigned int e[8];
#pragma unroll
      for(int u = 0; u < 8; u++)
         e[u] = k + u < m ? __builtin_nontemporal_load(&row[(k + u) * 64]) : 0u;
#pragma unroll
      for(int u = 0; u < 8; u++)
         if(k + u < m && e[u])
         {
            s += X[e[u] >> 4] * (float)(e[u] & 15u);
            cs += (int)(e[u] & 15u);
         }
   }
   const int d = cs / 3;                       // c_vv x_v, c_vv = deg v
   s += p * (float)d;
   if(!rf && d > 0)
   {
      s *= 0.25f / (float)d;
      p = p + (s - p) * omega;
   }
   __builtin_nontemporal_store(p, &xn[v]);
}


static bool MeshReady(GmlCtx *g)
{
   return !g->host_only && g->TypIdx[GmlVertices] && g->TypIdx[GmlTetrahedra];
}

} // namespace gml

using namespace gml;

extern "C" {

int GmlxTetQuality(size_t GmlIdx, int QalIdx)
{
   GML_CTX(g, GmlIdx);
   if(!MeshReady(g)) return -1;
   if(QalIdx < 1 || QalIdx > GmlMaxDat || !g->dat[QalIdx].use || g->dat[QalIdx].ItmTyp != GmlFlt
   || g->dat[QalIdx].NmbItm != 1 || g->dat[QalIdx].NmbLin != g->NmbEle[GmlTetrahedra])
      return -2;
   long long nt = g->NmbEle[GmlTetrahedra];
   const Shape sh = g->shape[GmlxBuiltinQuality];
   const v4i *tet = (const v4i *)g->dat[g->TypIdx[GmlTetrahedra]].dev;
   const v4f *crd = (const v4f *)g->dat[g->TypIdx[GmlVertices]].dev;
   float *q = (float *)g->dat[QalIdx].dev;
   g->dat[QalIdx].host_valid = false;
   BumpVersion(&g->dat[QalIdx]);
   if(!nt) return 1;
   int per = sh.per;
   if(per == 0)
   {
      // tile-staged form, when the tet table has a tile plan (built on the
      // second call for the same table)
      DownPlan *p = RepeatDownPlan(g, GmlTetrahedra);
      if(p && p->nrows == nt)
      {
         // R = tile rows / group size, 1 to 8
         int grp = std::min(std::min(sh.grp, 256), p->rows);
         if(p->rows / grp > 8) grp = p->rows / 8;
         const int R = p->rows / grp;
         auto kern = R == 8 ? k_tet_quality_tile<8> : R == 4 ? k_tet_quality_tile<4> : R == 2 ? k_tet_quality_tile<2>
                                                                                           : k_tet_quality_tile<1>;
         return TimedLaunch(g, GmlxBuiltinQuality, [&]() {
            hipLaunchKernelGGL(kern, dim3((unsigned int)p->ntile), dim3(grp), 0, g->stream, nt, p->tofs, p->tlist,
                               p->lid, crd, q, sh.remap);
         });
      }
      per = 2;
   }
   auto kern = per == 1 ? k_tet_quality<1> : per == 4 ? k_tet_quality<4> : k_tet_quality<2>;
   const long long span = (long long)(per == 1 || per == 4 ? per : 2) * sh.grp;
   return TimedLaunch(g, GmlxBuiltinQuality, [&]() {
      hipLaunchKernelGGL(kern, dim3((unsigned int)((nt + span - 1) / span)), dim3(sh.grp), 0, g->stream, nt, tet, crd,
                         q, sh.remap);
   });
}

int GmlxBallMin(size_t GmlIdx, int TetValIdx, int VerOutIdx)
{
   GML_CTX(g, GmlIdx);
   if(!MeshReady(g)) return -1;
   Data *a = (TetValIdx >= 1 && TetValIdx <= GmlMaxDat) ? &g->dat[TetValIdx] : nullptr;
   Data *o = (VerOutIdx >= 1 && VerOutIdx <= GmlMaxDat) ? &g->dat[VerOutIdx] : nullptr;
   if(!a || !o || !a->use || !o->use || a->ItmTyp != GmlFlt || o->ItmTyp != GmlFlt || a->NmbItm != 1 || o->NmbItm != 1
   || a->NmbLin != g->NmbEle[GmlTetrahedra] || o->NmbLin != g->NmbEle[GmlVertices])
      return -2;
   Uplink *u = CurrentUplink(g, GmlVertices, GmlTetrahedra);
   if(!u) return -3;
   long long nv = g->NmbEle[GmlVertices];
   const Shape sh = g->shape[GmlxBuiltinBallMin];
   o->host_valid = false;
   BumpVersion(o);
   if(!nv) return 1;
   return TimedLaunch(g, GmlxBuiltinBallMin, [&]() {
      hipLaunchKernelGGL(k_ball_min, dim3((unsigned int)((nv + sh.grp - 1) / sh.grp)), dim3(sh.grp), 0, g->stream,
                         nv, (const int *)g->dat[u->DegIdx].dev, (const unsigned int *)g->dat[u->OffIdx].dev,
                         (const int *)g->dat[u->BalIdx].dev, (const float *)a->dev, (float *)o->dev, sh.remap);
   });
}

int GmlxSmoothSweeps(size_t GmlIdx, int BaryIdx, int nsweep, double omega)
{
   GML_CTX(g, GmlIdx);
   if(!MeshReady(g) || nsweep < 0) return -1;
   Data *b = (BaryIdx >= 1 && BaryIdx <= GmlMaxDat) ? &g->dat[BaryIdx] : nullptr;
   if(!b || !b->use || b->ItmTyp != GmlFlt4 || b->NmbItm != 1 || b->NmbLin != g->NmbEle[GmlTetrahedra])
      return -2;
   Uplink *u = CurrentUplink(g, GmlVertices, GmlTetrahedra);
   if(!u) return -3;
   long long nv = g->NmbEle[GmlVertices], nt = g->NmbEle[GmlTetrahedra];
   // the alternate coordinate buffer follows the vertex count
   if(g->xalt && g->xalt_n != nv)
   {
      DevFree(g, g->xalt);
      g->xalt = nullptr;
   }
   Data *vd = &g->dat[g->TypIdx[GmlVertices]];
   const v4i *tet = (const v4i *)g->dat[g->TypIdx[GmlTetrahedra]].dev;
   const int *ref = (const int *)g->dat[g->RefIdx[GmlVertices]].dev;
   vd->host_valid = false;
   BumpVersion(vd);
   if(!nv || !nsweep) return 1;

   if(g->smooth_mode == 2)
   {
      const Shape sf = g->shape[GmlxBuiltinSweep];
      if(g->oplan && g->oplan->TV != sf.grp) { FreeSweepPlan(g, g->oplan); g->oplan = nullptr; }
      if(!g->oplan && !(g->oplan = NewSweepPlan(g, sf.grp, 1))) return -4;
      if(!g->xalt && !(g->xalt = DevAlloc(g, (size_t)nv * 16))) return -5;
      g->xalt_n = nv;
      SweepPlan *p = g->oplan;
      size_t lds = (size_t)p->max_nv * 16;
      if(lds > 160 * 1024) return -4;
      if(lds > 64 * 1024)
         HIPCHK(hipFuncSetAttribute((const void *)k_sweep_op, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      for(int i = 0; i < nsweep; i++)
      {
         const v4f *xo = (const v4f *)vd->dev;
         v4f *xn = (v4f *)g->xalt;
         int rc = TimedLaunch(g, GmlxBuiltinSweep, [&]() {
            hipLaunchKernelGGL(k_sweep_op, dim3((unsigned int)p->ntile), dim3(p->TV), lds, g->stream, nv,
                               p->vofs, p->vlist, p->noff, p->nbr, ref, xo, xn, (float)omega, sf.remap);
         });
         if(rc != 1) return rc;
         void *t = vd->dev; vd->dev = g->xalt; g->xalt = t;
      }
      return 1;
   }

   if(g->smooth_mode == 1)
   {
      const Shape sf = g->shape[GmlxBuiltinSweep];
      if(sf.grp > 256) return -4;              // the exact fused kernel takes tiles of up to 256
      if(g->plan && g->plan->TV != sf.grp) { FreeSweepPlan(g, g->plan); g->plan = nullptr; }
      if(!g->plan && !(g->plan = NewSweepPlan(g, sf.grp, 0))) return -4;
      if(!g->xalt && !(g->xalt = DevAlloc(g, (size_t)nv * 16))) return -5;
      g->xalt_n = nv;
      SweepPlan *p = g->plan;
      size_t lds = (size_t)(p->max_nv + p->max_nt) * 16;
      if(lds > 160 * 1024) return -4;
      if(lds > 64 * 1024)
         HIPCHK(hipFuncSetAttribute((const void *)k_sweep_fused, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      for(int i = 0; i < nsweep; i++)
      {
         const v4f *xo = (const v4f *)vd->dev;
         v4f *xn = (v4f *)g->xalt;
         int rc = TimedLaunch(g, GmlxBuiltinSweep, [&]() {
            hipLaunchKernelGGL(k_sweep_fused, dim3((unsigned int)p->ntile), dim3(p->TV), lds, g->stream, nv,
                               p->vofs, p->vlist, p->tofs, (const v4u16 *)p->tloc,
                               (const int *)g->dat[u->DegIdx].dev, (const unsigned int *)g->dat[u->OffIdx].dev,
                               p->lidx, ref, xo, xn, (float)omega, sf.remap);
         });
         if(rc != 1) return rc;
         // the new coordinates become the vertex table
         void *t = vd->dev; vd->dev = g->xalt; g->xalt = t;
      }
      return 1;
   }

   v4f *crd = (v4f *)vd->dev;
   const Shape ss = g->shape[GmlxBuiltinScatter], sg = g->shape[GmlxBuiltinGather];
   b->host_valid = false;
   BumpVersion(b);
   for(int i = 0; i < nsweep; i++)
   {
      int rc = TimedLaunch(g, GmlxBuiltinScatter, [&]() {
         hipLaunchKernelGGL(k_smooth_scatter, dim3((unsigned int)((nt + ss.grp - 1) / ss.grp)), dim3(ss.grp), 0,
                            g->stream, nt, tet, (const v4f *)crd, (v4f *)b->dev, ss.remap);
      });
      if(rc != 1) return rc;
      rc = TimedLaunch(g, GmlxBuiltinGather, [&]() {
         hipLaunchKernelGGL(k_smooth_gather, dim3((unsigned int)((nv + sg.grp - 1) / sg.grp)), dim3(sg.grp), 0,
                            g->stream, nv, (const int *)g->dat[u->DegIdx].dev, (const unsigned int *)g->dat[u->OffIdx].dev,
                            (const int *)g->dat[u->BalIdx].dev, (const v4f *)b->dev, ref, crd, (float)omega, sg.remap);
      });
      if(rc != 1) return rc;
   }
   return 1;
}

int GmlxSetSmoothMode(size_t GmlIdx, int mode)
{
   GML_CTX(g, GmlIdx);
   if(mode < 0 || mode > 2) return 0;
   g->smooth_mode = mode;
   return 1;
}

int GmlxGetSweepPlanInfo(size_t GmlIdx, long long *info)
{
   GML_CTX(g, GmlIdx);
   SweepPlan *p = g->smooth_mode == 2 ? g->oplan : g->plan;
   if(!p || !info) return 0;
   info[0] = p->TV; info[1] = p->ntile; info[2] = p->sum_nv; info[3] = p->sum_nt;
   info[4] = p->max_nv; info[5] = p->max_nt; info[6] = (long long)p->bytes;
   info[7] = p->sum_ring; info[8] = g->smooth_mode;
   return 1;
}

int GmlxSetBuiltinItems(size_t GmlIdx, int which, int PerThread)
{
   GML_CTX(g, GmlIdx);
   if(which != GmlxBuiltinQuality || (PerThread != 0 && PerThread != 1 && PerThread != 2 && PerThread != 4)) return 0;
   g->shape[which].per = PerThread;
   return 1;
}

int GmlxSetBuiltinShape(size_t GmlIdx, int which, int GrpSiz, int Remap)
{
   GML_CTX(g, GmlIdx);
   if(which < 0 || which > GmlxBuiltinCgUpdate || GrpSiz < 64 || GrpSiz > 1024 || (GrpSiz & 63)) return 0;
   if(which > GmlxBuiltinSweep && GrpSiz > 256) return 0;
   g->shape[which].grp = GrpSiz;
   g->shape[which].remap = Remap ? 1 : 0;
   return 1;
}

double GmlxGetBuiltinRunTime(size_t GmlIdx, int which, long long *launches)
{
   GML_CTX(g, GmlIdx);
   if(which < 0 || which >= kNmbBuiltin || g->host_only) return -1.;
   HIPCHK(hipStreamSynchronize(g->stream));
   KernStat *st = Stat(g, which);
   HarvestEvents(st->ev, &st->time);
   if(launches) *launches = st->launches;
   return st->time;
}

} // extern "C"
