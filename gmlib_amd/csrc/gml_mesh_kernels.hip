// Hand-written HIP kernels for the mesh workloads of this engine: tet mean-
// ratio quality, per-vertex ball minimum, and Jacobi barycentre smoothing
// (tet scatter + vertex-ball gather).  They compute exactly what the
// generated kernels built from gmlib_amd/kernels/*.cl compute, expression by
// expression, so their results are bit-identical to the drop-in path; what
// differs is how they move data:
//   - element rows are streamed with non-temporal loads (read once),
//   - the vertex ball is read from its SELL-64 slabs with one 256-byte load
//     per slot and wave, and the barycentre gathers of a row are issued in
//     batches of 8 before any is consumed,
//   - blocks are remapped so that each XCD (own 4 MB L2) walks a contiguous
//     range of the Hilbert-ordered mesh.
// Reference patterns: examples/MeshQuality (quality), examples/BallQuality
// (ball min), examples/NodeSmoothing (scatter + gather).

#include "gml_internal.h"

namespace gml {

struct KernStat {
   std::vector<LaunchRec> ev;
   double time = 0.;
   long long launches = 0;
};

// per-handle timing records of the built-in kernels (index = GmlxBuiltin*)
static std::mutex g_stat_mtx;
static std::vector<std::pair<GmlCtx *, KernStat *>> g_stats;

static KernStat *Stat(GmlCtx *g, int which)
{
   std::lock_guard<std::mutex> lk(g_stat_mtx);
   for(auto &p : g_stats)
      if(p.first == g) return &p.second[which];
   KernStat *s = new KernStat[8];
   g_stats.push_back({g, s});
   return &s[which];
}

void FreeBuiltinStats(GmlCtx *g)
{
   std::lock_guard<std::mutex> lk(g_stat_mtx);
   for(size_t i = 0; i < g_stats.size(); i++)
      if(g_stats[i].first == g)
      {
         for(int w = 0; w < 8; w++)
            for(auto &e : g_stats[i].second[w].ev) { (void)hipEventDestroy(e.beg); (void)hipEventDestroy(e.end); }
         delete[] g_stats[i].second;
         g_stats.erase(g_stats.begin() + i);
         return;
      }
}

__device__ __forceinline__ unsigned int xcd_block(int remap)
{
   unsigned int b = blockIdx.x;
   if(remap)
   {
      unsigned int nb = gridDim.x, q = nb >> 3, r = nb & 7, x = b & 7, i = b >> 3;
      b = (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
   }
   return b;
}

__device__ __forceinline__ int4 ld_nt(const int4 *p)
{
   int4 v;
   v.x = __builtin_nontemporal_load(&p->x);
   v.y = __builtin_nontemporal_load(&p->y);
   v.z = __builtin_nontemporal_load(&p->z);
   v.w = __builtin_nontemporal_load(&p->w);
   return v;
}

__device__ __forceinline__ float dot4(float4 a, float4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }

// mean ratio: the expression sequence of tet_meanratio.cl
__device__ __forceinline__ float meanratio(float4 p0, float4 p1, float4 p2, float4 p3)
{
   float4 e0 = p1 - p0, e1 = p2 - p0, e2 = p3 - p0, e3 = p2 - p1, e4 = p3 - p1, e5 = p3 - p2;
   float s = dot4(e0, e0) + dot4(e1, e1) + dot4(e2, e2) + dot4(e3, e3) + dot4(e4, e4) + dot4(e5, e5);
   float4 c = make_float4(e0.y * e1.z - e0.z * e1.y, e0.z * e1.x - e0.x * e1.z, e0.x * e1.y - e0.y * e1.x, 0.f);
   float v = dot4(c, e2) * (1.f / 6.f);
   float q = 12.f * cbrtf(9.f * v * v) / s;
   if(v < 0.f) q = -q;
   return q;
}

__global__ void __launch_bounds__(1024) k_tet_quality(long long nt, const int4 *__restrict__ tet,
                                                       const float4 *__restrict__ crd, float *__restrict__ q, int remap)
{
   long long t = (long long)xcd_block(remap) * blockDim.x + threadIdx.x;
   if(t >= nt) return;
   int4 v = ld_nt(&tet[t]);
   q[t] = meanratio(crd[v.x], crd[v.y], crd[v.z], crd[v.w]);
}

__global__ void __launch_bounds__(1024) k_smooth_scatter(long long nt, const int4 *__restrict__ tet,
                                                          const float4 *__restrict__ crd, float4 *__restrict__ bary, int remap)
{
   long long t = (long long)xcd_block(remap) * blockDim.x + threadIdx.x;
   if(t >= nt) return;
   int4 v = ld_nt(&tet[t]);
   float4 a = crd[v.x], b = crd[v.y], c = crd[v.z], d = crd[v.w];
   bary[t] = (((a + b) + c) + d) * 0.25f;
}

// Ball gather: slot k of vertex v is ent[(off[v >> 6] + k) * 64 + (v & 63)].
__global__ void __launch_bounds__(1024) k_smooth_gather(long long nv, const int *__restrict__ deg,
                                                         const unsigned int *__restrict__ off, const int *__restrict__ ent,
                                                         const float4 *__restrict__ bary, const int *__restrict__ ref,
                                                         float4 *__restrict__ crd, float omega, int remap)
{
   long long v = (long long)xcd_block(remap) * blockDim.x + threadIdx.x;
   if(v >= nv) return;
   const int d = deg[v];
   const int *row = ent + (size_t)off[v >> 6] * 64 + (v & 63);
   float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
   int k = 0;
   for(; k + 8 <= d; k += 8)
   {
      int e[8];
      float4 b[8];
#pragma unroll
      for(int i = 0; i < 8; i++) e[i] = __builtin_nontemporal_load(&row[(k + i) * 64]) >> 4;
#pragma unroll
      for(int i = 0; i < 8; i++) b[i] = bary[e[i]];
#pragma unroll
      for(int i = 0; i < 8; i++) s += b[i];
   }
   for(; k < d; k++)
      s += bary[__builtin_nontemporal_load(&row[k * 64]) >> 4];
   const int r = ref[v];
   if(!r && d > 0)
   {
      float4 p = crd[v];
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
   float m = d > 0 ? val[row[0] >> 4] : 0.f;
   for(int k = 1; k < d; k++)
      m = fminf(m, val[row[k * 64] >> 4]);
   out[v] = m;
}

// Fused Jacobi sweep over one tile of TV = blockDim.x vertices (see the
// tile plan, gml_plan.cpp): stage the tile's vertex coordinates, form its
// distinct barycentres, gather every ball from LDS, write the new
// coordinates to the other buffer.
__global__ void __launch_bounds__(256) k_sweep_fused(long long nv, const unsigned int *__restrict__ vofs,
                                                      const int *__restrict__ vlist, const unsigned int *__restrict__ tofs,
                                                      const ushort4 *__restrict__ tloc, const int *__restrict__ deg,
                                                      const unsigned int *__restrict__ off, const unsigned short *__restrict__ lidx,
                                                      const int *__restrict__ ref, const float4 *__restrict__ xo,
                                                      float4 *__restrict__ xn, float omega, int remap)
{
   extern __shared__ float4 lds[];
   const unsigned int tile = xcd_block(remap);
   const int tid = threadIdx.x, TV = blockDim.x;
   const unsigned int v0 = vofs[tile], nvt = vofs[tile + 1] - v0;
   const unsigned int t0 = tofs[tile], ntt = tofs[tile + 1] - t0;
   float4 *X = lds, *Bc = lds + nvt;
   for(unsigned int i = tid; i < nvt; i += TV)
      X[i] = xo[__builtin_nontemporal_load(&vlist[v0 + i])];
   __syncthreads();
   for(unsigned int j = tid; j < ntt; j += TV)
   {
      ushort4 l = tloc[t0 + j];
      Bc[j] = (((X[l.x] + X[l.y]) + X[l.z]) + X[l.w]) * 0.25f;
   }
   __syncthreads();
   const long long v = (long long)tile * TV + tid;
   if(v >= nv) return;
   const int d = deg[v];
   const unsigned short *row = lidx + (size_t)off[v >> 6] * 64 + (v & 63);
   float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
   for(int k = 0; k < d; k++)
      s += Bc[__builtin_nontemporal_load(&row[k * 64])];
   float4 p = xo[v];
   if(!ref[v] && d > 0)
   {
      s *= 1.f / (float)d;
      p = p + (s - p) * omega;
   }
   xn[v] = p;
}

struct Shape { int grp = 256, remap = 1; };
static Shape g_shape[8];
static int g_smooth_fused = 1;

template<class F>
static int TimedLaunch(GmlCtx *g, int which, F launch)
{
   KernStat *st = Stat(g, which);
   LaunchRec r;
   bool ev = hipEventCreate(&r.beg) == hipSuccess && hipEventCreate(&r.end) == hipSuccess;
   if(ev) (void)hipEventRecord(r.beg, g->stream);
   launch();
   if(hipGetLastError() != hipSuccess) return -6;
   if(ev)
   {
      (void)hipEventRecord(r.end, g->stream);
      st->ev.push_back(r);
      if(st->ev.size() > 256) HarvestEvents(st->ev, &st->time);
   }
   st->launches++;
   return 1;
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
   const Shape sh = g_shape[GmlxBuiltinQuality];
   const int4 *tet = (const int4 *)g->dat[g->TypIdx[GmlTetrahedra]].dev;
   const float4 *crd = (const float4 *)g->dat[g->TypIdx[GmlVertices]].dev;
   float *q = (float *)g->dat[QalIdx].dev;
   g->dat[QalIdx].host_valid = false;
   if(!nt) return 1;
   return TimedLaunch(g, GmlxBuiltinQuality, [&]() {
      hipLaunchKernelGGL(k_tet_quality, dim3((unsigned int)((nt + sh.grp - 1) / sh.grp)), dim3(sh.grp), 0, g->stream,
                         nt, tet, crd, q, sh.remap);
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
   if(!g->up[GmlVertices][GmlTetrahedra].BalIdx && !BuildUplink(g, GmlVertices, GmlTetrahedra)) return -3;
   Uplink *u = &g->up[GmlVertices][GmlTetrahedra];
   long long nv = g->NmbEle[GmlVertices];
   const Shape sh = g_shape[GmlxBuiltinBallMin];
   o->host_valid = false;
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
   if(!g->up[GmlVertices][GmlTetrahedra].BalIdx && !BuildUplink(g, GmlVertices, GmlTetrahedra)) return -3;
   Uplink *u = &g->up[GmlVertices][GmlTetrahedra];
   long long nv = g->NmbEle[GmlVertices], nt = g->NmbEle[GmlTetrahedra];
   Data *vd = &g->dat[g->TypIdx[GmlVertices]];
   const int4 *tet = (const int4 *)g->dat[g->TypIdx[GmlTetrahedra]].dev;
   const int *ref = (const int *)g->dat[g->RefIdx[GmlVertices]].dev;
   vd->host_valid = false;
   if(!nv || !nsweep) return 1;

   if(g_smooth_fused)
   {
      const Shape sf = g_shape[GmlxBuiltinSweep];
      if(g->plan && g->plan->TV != sf.grp) { FreeSweepPlan(g, g->plan); g->plan = nullptr; }
      if(!g->plan && !(g->plan = NewSweepPlan(g, sf.grp))) return -4;
      if(!g->xalt && !(g->xalt = DevAlloc(g, (size_t)nv * 16))) return -5;
      SweepPlan *p = g->plan;
      size_t lds = (size_t)(p->max_nv + p->max_nt) * 16;
      if(lds > 160 * 1024) return -4;
      if(lds > 64 * 1024)
         HIPCHK(hipFuncSetAttribute((const void *)k_sweep_fused, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      for(int i = 0; i < nsweep; i++)
      {
         const float4 *xo = (const float4 *)vd->dev;
         float4 *xn = (float4 *)g->xalt;
         int rc = TimedLaunch(g, GmlxBuiltinSweep, [&]() {
            hipLaunchKernelGGL(k_sweep_fused, dim3((unsigned int)p->ntile), dim3(p->TV), lds, g->stream, nv,
                               p->vofs, p->vlist, p->tofs, (const ushort4 *)p->tloc,
                               (const int *)g->dat[u->DegIdx].dev, (const unsigned int *)g->dat[u->OffIdx].dev,
                               p->lidx, ref, xo, xn, (float)omega, sf.remap);
         });
         if(rc != 1) return rc;
         // the new coordinates become the vertex table
         void *t = vd->dev; vd->dev = g->xalt; g->xalt = t;
      }
      return 1;
   }

   float4 *crd = (float4 *)vd->dev;
   const Shape ss = g_shape[GmlxBuiltinScatter], sg = g_shape[GmlxBuiltinGather];
   b->host_valid = false;
   for(int i = 0; i < nsweep; i++)
   {
      int rc = TimedLaunch(g, GmlxBuiltinScatter, [&]() {
         hipLaunchKernelGGL(k_smooth_scatter, dim3((unsigned int)((nt + ss.grp - 1) / ss.grp)), dim3(ss.grp), 0,
                            g->stream, nt, tet, (const float4 *)crd, (float4 *)b->dev, ss.remap);
      });
      if(rc != 1) return rc;
      rc = TimedLaunch(g, GmlxBuiltinGather, [&]() {
         hipLaunchKernelGGL(k_smooth_gather, dim3((unsigned int)((nv + sg.grp - 1) / sg.grp)), dim3(sg.grp), 0,
                            g->stream, nv, (const int *)g->dat[u->DegIdx].dev, (const unsigned int *)g->dat[u->OffIdx].dev,
                            (const int *)g->dat[u->BalIdx].dev, (const float4 *)b->dev, ref, crd, (float)omega, sg.remap);
      });
      if(rc != 1) return rc;
   }
   return 1;
}

int GmlxSetSmoothMode(int fused)
{
   g_smooth_fused = fused ? 1 : 0;
   return 1;
}

int GmlxGetSweepPlanInfo(size_t GmlIdx, long long *info)
{
   GML_CTX(g, GmlIdx);
   SweepPlan *p = g->plan;
   if(!p || !info) return 0;
   info[0] = p->TV; info[1] = p->ntile; info[2] = p->sum_nv; info[3] = p->sum_nt;
   info[4] = p->max_nv; info[5] = p->max_nt; info[6] = (long long)p->bytes;
   return 1;
}

int GmlxSetBuiltinShape(int which, int GrpSiz, int Remap)
{
   if(which < 0 || which >= 8 || GrpSiz < 64 || GrpSiz > 1024 || (GrpSiz & 63)) return 0;
   if(which == GmlxBuiltinSweep && GrpSiz > 256) return 0;
   g_shape[which].grp = GrpSiz;
   g_shape[which].remap = Remap ? 1 : 0;
   return 1;
}

double GmlxGetBuiltinRunTime(size_t GmlIdx, int which, long long *launches)
{
   GML_CTX(g, GmlIdx);
   if(which < 0 || which >= 8 || g->host_only) return -1.;
   HIPCHK(hipStreamSynchronize(g->stream));
   KernStat *st = Stat(g, which);
   HarvestEvents(st->ev, &st->time);
   if(launches) *launches = st->launches;
   return st->time;
}

} // extern "C"
