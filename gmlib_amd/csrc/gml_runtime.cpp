// Runtime layer of libgml: handle, data registry, host<->device transfers,
// metrics and info queries of the GMlib C API over HIP.
//
// Reference behaviour followed (paths in the GMlib tree):
//   GmlInit/GmlStop/GmlListGPU            sources/gmlib3.c:385-518
//   GmlNewParameters/MeshData/Solution/Link sources/gmlib3.c:525-736
//   NewData/GmlFreeData                   sources/gmlib3.c:1779-1834
//   GmlSet/GetDataLine, Set/GetDataBlock  sources/gmlib3.c:1841-2105
//   Upload/Download, parameters           sources/gmlib3.c:2112-2200
//   metrics, info, numbering              sources/gmlib3.c:3648-3712, 4274-4453

#include "gml_internal.h"

#include <atomic>

#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <sys/time.h>

namespace gml {

// ---------------------------------------------------------------------------
// Interface tables.  Their values are fixed by the reference interface
// (sizes of the OpenCL types, the names user bodies see, link widths).
// ---------------------------------------------------------------------------
const int OclTypSiz[GmlMaxOclTyp] = {4, 8, 16, 32, 64, 4, 8, 16, 32, 64,
                                     8, 16, 32, 64, 128, 1, 2, 4, 8, 16};
const int TypVecSiz[GmlMaxOclTyp] = {1, 2, 4, 8, 16, 1, 2, 4, 8, 16,
                                     1, 2, 4, 8, 16, 1, 2, 4, 8, 16};
const char *OclTypStr[GmlMaxOclTyp] = {
   "int", "int2", "int4", "int8", "int16", "float", "float2", "float4", "float8", "float16",
   "double", "double2", "double4", "double8", "double16", "char", "char2", "char4", "char8", "char16"};
const char *OclNulVec[GmlMaxOclTyp] = {
   "(int)0", "(int2)(0)", "(int4)(0)", "(int8)(0)", "(int16)(0)",
   "0.f", "(float2)(0.f)", "(float4)(0.f)", "(float8)(0.f)", "(float16)(0.f)",
   "0.", "(double2)(0.)", "(double4)(0.)", "(double8)(0.)", "(double16)(0.)",
   "(char)0", "(char2)(0)", "(char4)(0)", "(char8)(0)", "(char16)(0)"};
const int MshItmTyp[GmlMaxEleTyp] = {GmlFlt4, GmlInt2, GmlInt4, GmlInt4, GmlInt4, GmlInt8, GmlInt8, GmlInt8};
const int EleNmbNod[GmlMaxEleTyp] = {0, 2, 3, 4, 4, 5, 6, 8};
const int MshTypDim[GmlMaxEleTyp] = {0, 1, 2, 2, 3, 3, 3, 3};
const char *BalTypStr[GmlMaxEleTyp] = {"Ver", "Edg", "Tri", "Qad", "Tet", "Pyr", "Pri", "Hex"};
const char *MshTypStr[GmlMaxEleTyp] = {"VerCrd", "EdgVer", "TriVer", "QadVer", "TetVer", "PyrVer", "PriVer", "HexVer"};
const char *MshRefStr[GmlMaxEleTyp] = {"VerRef", "EdgRef", "TriRef", "QadRef", "TetRef", "PyrRef", "PriRef", "HexRef"};

// Base and maximum link widths (source row x destination column).
const int LenMatBas[GmlMaxEleTyp][GmlMaxEleTyp] = {
   {0,16, 8, 4,32,16,16, 8}, {2, 0, 2, 2, 8, 8, 8, 4}, {4, 4, 0, 0, 2, 2, 2, 0}, {4, 4, 0, 0, 0, 2, 2, 2},
   {4, 8, 4, 0, 4, 4, 4, 0}, {8, 8, 4, 1, 4, 8, 8, 1}, {8,16, 2, 4, 2, 8, 8, 4}, {8,16, 0, 8, 0, 8, 8, 8}};
const int LenMatMax[GmlMaxEleTyp][GmlMaxEleTyp] = {
   {0,64,16, 8,64,32,32,16}, {2, 0, 2, 2,16,16,16, 8}, {4, 4, 0, 0, 2, 2, 2, 0}, {4, 4, 0, 0, 0, 2, 2, 2},
   {4, 8, 4, 0, 4, 4, 4, 0}, {8, 8, 4, 1, 4, 8, 8, 1}, {8,16, 2, 4, 2, 8, 8, 4}, {8,16, 0, 8, 0, 8, 8, 8}};
// Number of sub-entities of each kind per element (row = element kind).
const int NmbTpoLnk[GmlMaxEleTyp][GmlMaxEleTyp] = {
   {1, 1, 1, 1, 1, 1, 1, 1}, {2, 2, 1, 1, 1, 1, 1, 1}, {3, 3, 3, 0, 2, 2, 2, 0}, {4, 4, 0, 4, 0, 2, 2, 2},
   {4, 6, 4, 0, 4, 4, 4, 0}, {5, 8, 4, 1, 4, 5, 5, 1}, {6, 9, 2, 3, 2, 5, 5, 3}, {8,12, 0, 6, 0, 6, 6, 6}};
// Local edge and face numbering of each element kind: the local index a
// ball entry carries (its "voyeur"), so it must match the reference's.
const int ItmEdgNod[8][12][2] = {
 {{0,0}},
 {{0,1}},
 {{0,1},{1,2},{2,0}},
 {{0,1},{1,2},{2,3},{3,0}},
 {{0,1},{1,2},{2,0},{3,0},{3,1},{3,2}},
 {{0,1},{3,2},{0,3},{1,2},{0,4},{1,4},{2,4},{3,4}},
 {{0,1},{1,2},{2,0},{3,4},{4,5},{5,3},{0,3},{1,4},{2,5}},
 {{0,1},{3,2},{7,6},{4,5},{0,3},{4,7},{5,6},{1,2},{0,4},{1,5},{2,6},{3,7}}};
const int ItmFacDeg[8][6] = {
 {0,0,0,0,0,0}, {0,0,0,0,0,0}, {3,0,0,0,0,0}, {4,0,0,0,0,0},
 {3,3,3,3,0,0}, {3,3,3,3,4,0}, {3,3,4,4,4,0}, {4,4,4,4,4,4}};
const int ItmFacNod[8][6][4] = {
 {{0}}, {{0}},
 {{0,1,2,0}},
 {{0,1,2,3}},
 {{1,2,3,0},{2,0,3,0},{3,0,1,0},{0,2,1,0}},
 {{0,1,4,0},{1,2,4,0},{2,3,4,0},{3,0,4,0},{3,2,1,0}},
 {{0,2,1,0},{3,4,5,0},{0,1,4,3},{1,2,5,4},{3,5,2,0}},
 {{0,4,7,3},{1,2,6,5},{0,1,5,4},{3,7,6,2},{0,3,2,1},{4,5,6,7}}};

void DbgPrint(GmlCtx *g, const char *fmt, ...)
{
   if(!g || !g->DbgFlg) return;
   va_list ap;
   va_start(ap, fmt);
   vfprintf(stdout, fmt, ap);
   va_end(ap);
   fflush(stdout);
}

int NewDatIdx(GmlCtx *g)
{
   for(int i = 1; i <= GmlMaxDat; i++)
      if(!g->dat[i].use)
      {
         g->dat[i] = Data();
         g->dat[i].use = true;
         return i;
      }
   fprintf(stderr, "gml: all %d data slots are in use\n", GmlMaxDat);
   return 0;
}

// Allocate the device buffer of a data item (zero-filled).  The host mirror
// is allocated lazily by EnsureHost.
bool AllocData(GmlCtx *g, Data *d, bool zero)
{
   size_t siz = d->MemSiz ? d->MemSiz : 16;
   if(g->budget && g->MemSiz + siz > g->budget)
   {
      fprintf(stderr, "gml: allocation of %zu MB exceeds the device budget (%zu MB used)\n",
              siz >> 20, g->MemSiz >> 20);
      d->use = false;
      return false;
   }
   if(g->host_only)
      d->dev = calloc(1, siz);
   else if(hipMalloc(&d->dev, siz) != hipSuccess)
      d->dev = nullptr;
   if(!d->dev)
   {
      (void)hipGetLastError();
      fprintf(stderr, "gml: cannot allocate %zu MB on the GPU (%zu MB already used)\n",
              siz >> 20, g->MemSiz >> 20);
      d->dev = nullptr;
      d->use = false;
      return false;
   }
   if(zero && !g->host_only)
      HIPCHK(hipMemsetAsync(d->dev, 0, siz, g->stream));
   g->MemSiz += d->MemSiz;
   BumpVersion(d);
   return true;
}

bool EnsureHost(Data *d)
{
   if(d->host) return true;
   d->host = calloc(1, d->MemSiz ? d->MemSiz : 16);
   if(!d->host)
   {
      fprintf(stderr, "gml: cannot allocate %zu MB on the CPU\n", d->MemSiz >> 20);
      return false;
   }
   d->host_valid = false;
   return true;
}

void FreeDataSlot(GmlCtx *g, int idx)
{
   Data *d = &g->dat[idx];
   if(!d->use) return;
   if(d->dev)
   {
      DevFree(g, d->dev);
      g->MemSiz -= d->MemSiz;
   }
   free(d->host);
   *d = Data();
}

static std::atomic<unsigned long long> g_version{0};

void BumpVersion(Data *d) { d->ver = ++g_version; }

size_t Upload(GmlCtx *g, int idx)
{
   if(idx < 1 || idx > GmlMaxDat) return 0;
   Data *d = &g->dat[idx];
   if(!d->use || !d->dev || !d->host) return 0;
   if(!DevCopy(g, d->dev, d->host, d->MemSiz, 1))
   {
      fprintf(stderr, "gml: uploading data %d to the GPU failed\n", idx);
      return 0;
   }
   d->host_valid = true;
   BumpVersion(d);
   g->MovSiz += d->MemSiz;
   return d->MemSiz;
}

size_t Download(GmlCtx *g, int idx)
{
   if(idx < 1 || idx > GmlMaxDat) return 0;
   Data *d = &g->dat[idx];
   if(!d->use || !d->dev || !EnsureHost(d)) return 0;
   if(!DevCopy(g, d->host, d->dev, d->MemSiz, 2))
   {
      fprintf(stderr, "gml: downloading data %d from the GPU failed\n", idx);
      return 0;
   }
   d->host_valid = true;
   g->MovSiz += d->MemSiz;
   return d->MemSiz;
}

void *DevAlloc(GmlCtx *g, size_t bytes)
{
   void *p = nullptr;
   if(g->host_only) return calloc(1, bytes ? bytes : 16);
   if(hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) { (void)hipGetLastError(); return nullptr; }
   return p;
}

void DevFree(GmlCtx *g, void *p)
{
   if(!p) return;
   if(g->host_only) { free(p); return; }
   HIPCHK(hipStreamSynchronize(g->stream));
   HIPCHK(hipFree(p));
}

// Ordered with the handle's stream and synchronous for the caller.
bool DevCopy(GmlCtx *g, void *dst, const void *src, size_t bytes, int dir)
{
   if(g->host_only) { memcpy(dst, src, bytes); return true; }
   hipMemcpyKind k = dir == 1 ? hipMemcpyHostToDevice : (dir == 2 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
   return hipMemcpyAsync(dst, src, bytes, k, g->stream) == hipSuccess
       && hipStreamSynchronize(g->stream) == hipSuccess;
}

// Vector width rule of generated code (GetCntVec, gmlib3.c:2945-2964), without
// the 128-entry cap: p = ceil(log2 n); p <= 4 -> one vector of 2^p ints,
// otherwise 2^(p-4) vectors of 16.
int GetCntVec(int siz, int *cnt, int *vec, int *typ)
{
   static const int VecTyp[5] = {GmlInt, GmlInt2, GmlInt4, GmlInt8, GmlInt16};
   int p = 0;
   while((1 << p) < siz) p++;
   if(p > 4) { *cnt = 1 << (p - 4); *vec = 16; *typ = GmlInt16; }
   else      { *cnt = 1; *vec = 1 << p; *typ = VecTyp[p]; }
   return p;
}

// a mesh table and its refs (zero-filled unless the caller writes every row
// and ref itself); the element data index, 0 on failure
int NewMeshTable(GmlCtx *g, int MshTyp, int NmbLin, bool zero)
{
   if(MshTyp < 0 || MshTyp >= GmlMaxEleTyp || NmbLin < 0) return 0;
   int EleIdx = NewDatIdx(g);
   if(!EleIdx) return 0;
   Data *e = &g->dat[EleIdx];
   e->kind = DkEle;
   e->MshTyp = MshTyp;
   e->ItmTyp = MshItmTyp[MshTyp];
   e->NmbItm = 1;
   e->ItmSiz = OclTypSiz[e->ItmTyp];
   e->ItmLen = TypVecSiz[e->ItmTyp];
   e->NmbLin = NmbLin;
   e->LinSiz = (size_t)e->ItmSiz;
   e->MemSiz = (size_t)NmbLin * e->LinSiz;
   e->nam = MshTypStr[MshTyp];
   if(!AllocData(g, e, zero)) return 0;

   int RefIdx = NewDatIdx(g);
   if(!RefIdx)
   {
      FreeDataSlot(g, EleIdx);
      return 0;
   }
   Data *r = &g->dat[RefIdx];
   r->kind = DkRef;
   r->MshTyp = MshTyp;
   r->ItmTyp = GmlInt;
   r->NmbItm = 1;
   r->ItmSiz = 4;
   r->ItmLen = 0;
   r->NmbLin = NmbLin;
   r->LinSiz = 4;
   r->MemSiz = (size_t)NmbLin * 4;
   r->nam = MshRefStr[MshTyp];
   if(!AllocData(g, r, zero))
   {
      FreeDataSlot(g, EleIdx);
      return 0;
   }

   g->NmbEle[MshTyp] = NmbLin;
   g->TypIdx[MshTyp] = EleIdx;
   g->LnkMat[MshTyp][GmlVertices] = EleIdx;   // the element table is the downlink to vertices
   g->RefIdx[MshTyp] = RefIdx;
   DbgPrint(g, "Allocated mesh data: index=%2d, size=%zu bytes\nAllocated ref  data: index=%2d, size=%zu bytes\n",
            EleIdx, e->MemSiz, RefIdx, r->MemSiz);
   return EleIdx;
}

} // namespace gml

using namespace gml;

// ===========================================================================
// C API
// ===========================================================================
extern "C" {

size_t GmlInit(int DevIdx)
{
   int n = 0;
   if(hipGetDeviceCount(&n) != hipSuccess || n == 0)
   {
      (void)hipGetLastError();
      puts("Could not find a HIP device to run GMlib kernels.");
      return 0;
   }
   if(DevIdx < 0 || DevIdx >= n)
   {
      printf("Selected device Id is out of bounds (0 -> %d)\n", n - 1);
      return 0;
   }
   GmlCtx *g = new GmlCtx();
   g->dev = DevIdx;
   if(hipSetDevice(DevIdx) != hipSuccess
   || hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess)
   {
      puts("HIP stream creation failed.");
      delete g;
      return 0;
   }
   // FP64 is native on gfx950
   g->DblExt = 1;
   if(hipMalloc(&g->red_scratch, ReduceScratchBytes()) != hipSuccess
   || hipMemsetAsync(g->red_scratch, 0, ReduceScratchBytes(), g->stream) != hipSuccess
   || hipHostMalloc((void **)&g->red_host, 64, hipHostMallocDefault) != hipSuccess
   || hipMalloc(&g->dummy_par, 256) != hipSuccess
   || hipMemsetAsync(g->dummy_par, 0, 256, g->stream) != hipSuccess)
   {
      puts("GMlib device scratch allocation failed.");
      delete g;
      return 0;
   }
   // device address of the no-allocation reduction scratch (once per device)
   if(!NoAllocScratchInit(DevIdx))
      puts("GmlxReduceVectorNoAlloc is unavailable on this device.");
   HIPCHK(hipStreamSynchronize(g->stream));
   return reinterpret_cast<size_t>(g);
}

// A handle with no device: data live in host memory, links are built on the
// host, kernels are generated and compiled for gfx950 but launches and
// reductions return errors.  For inspection and compile checks on machines
// without a GPU.
size_t GmlxInitHostOnly(void)
{
   GmlCtx *g = new GmlCtx();
   g->host_only = true;
   g->dev = -1;
   return reinterpret_cast<size_t>(g);
}

void GmlStop(size_t GmlIdx)
{
   GML_CTX(g, GmlIdx);
   if(!g) return;
   if(g->host_only)
   {
      FreeAlgebra(g);
      for(int i = 1; i <= GmlMaxDat; i++)
         FreeDataSlot(g, i);
      delete g;
      return;
   }
   HIPCHK(hipStreamSynchronize(g->stream));
   for(int i = 1; i <= GmlMaxKrn; i++)
   {
      Kernel *k = &g->krn[i];
      for(auto &e : k->evts) FreeRec(e);
      for(auto &e : k->tune_ev) FreeRec(e);
      if(k->mod) (void)hipModuleUnload(k->mod);
      if(k->mod2) (void)hipModuleUnload(k->mod2);
      if(k->mod3) (void)hipModuleUnload(k->mod3);
   }
   for(int r = 0; r < GmlMaxRed; r++)
      for(auto &e : g->red_evts[r]) FreeRec(e);
   FreeBuiltinStats(g);
   FreeSweepPlan(g, g->plan);
   FreeSweepPlan(g, g->oplan);
   g->plan = g->oplan = nullptr;
   if(g->xalt) { DevFree(g, g->xalt); g->xalt = nullptr; }
   FreeBatches(g);
   for(int a = 0; a < GmlMaxEleTyp; a++)
      for(int b = 0; b < GmlMaxEleTyp; b++)
         for(auto &sp : g->up[a][b].stage) { FreeStagePlan(g, sp); sp = nullptr; }
   for(auto &dp : g->dplan) { FreeDownPlan(g, dp); dp = nullptr; }
   FreeAlgebra(g);
   for(int i = 1; i <= GmlMaxDat; i++)
      FreeDataSlot(g, i);
   (void)hipFree(g->red_scratch);
   (void)hipHostFree(g->red_host);
   (void)hipFree(g->dummy_par);
   FreeItemPool(g);
   (void)hipStreamDestroy(g->stream);
   delete g;
}

void GmlListGPU()
{
   int n = 0;
   if(hipGetDeviceCount(&n) != hipSuccess)
   {
      (void)hipGetLastError();
      puts("Getting the list of GPU failed");
      return;
   }
   for(int i = 0; i < n; i++)
   {
      hipDeviceProp_t p;
      if(hipGetDeviceProperties(&p, i) == hipSuccess)
         printf("      %d      : %s (%s)\n", i, p.name, p.gcnArchName);
   }
}

void *GmlNewParameters(size_t GmlIdx, int siz, char *src)
{
   GML_CTX(g, GmlIdx);
   int idx = NewDatIdx(g);
   if(!idx) return nullptr;
   Data *d = &g->dat[idx];
   d->kind = DkArg;
   d->NmbItm = 1;
   d->ItmSiz = siz;
   d->NmbLin = 1;
   d->LinSiz = (size_t)siz;
   d->MemSiz = (size_t)siz;
   d->nam = "GmlParSct";
   d->src = src;
   if(!AllocData(g, d) || !EnsureHost(d))
      return nullptr;
   g->ParIdx = idx;
   DbgPrint(g, "Allocated a parameters data: index=%2d, size=%zu bytes\n", idx, d->MemSiz);
   return d->host;
}

int GmlNewMeshData(size_t GmlIdx, int MshTyp, int NmbLin)
{
   GML_CTX(g, GmlIdx);
   return NewMeshTable(g, MshTyp, NmbLin, true);
}

int GmlNewSolutionData(size_t GmlIdx, int MshTyp, int NmbDat, int ItmTyp, char *nam)
{
   GML_CTX(g, GmlIdx);
   if(MshTyp < 0 || MshTyp >= GmlMaxEleTyp) return 0;
   if(ItmTyp < GmlInt || ItmTyp >= GmlMaxOclTyp || NmbDat < 1) return 0;
   int idx = NewDatIdx(g);
   if(!idx) return 0;
   Data *d = &g->dat[idx];
   d->kind = DkRaw;
   d->MshTyp = MshTyp;
   d->ItmTyp = ItmTyp;
   d->NmbItm = NmbDat;
   d->ItmSiz = OclTypSiz[ItmTyp];
   d->ItmLen = TypVecSiz[ItmTyp];
   d->NmbLin = g->NmbEle[MshTyp];
   d->LinSiz = (size_t)NmbDat * d->ItmSiz;
   d->MemSiz = (size_t)d->NmbLin * d->LinSiz;
   d->nam = nam ? nam : "";
   if(!AllocData(g, d)) return 0;
   DbgPrint(g, "Allocated solutions data: index=%2d, size=%zu bytes\n", idx, d->MemSiz);
   return idx;
}

int GmlNewLinkData(size_t GmlIdx, int MshTyp, int LnkTyp, int NmbDat, char *nam)
{
   GML_CTX(g, GmlIdx);
   if(MshTyp < 0 || MshTyp >= GmlMaxEleTyp || LnkTyp < 0 || LnkTyp >= GmlMaxEleTyp || NmbDat < 1) return 0;
   int idx = NewDatIdx(g);
   if(!idx) return 0;
   int cnt, vec, typ;
   GetCntVec(NmbDat, &cnt, &vec, &typ);
   Data *d = &g->dat[idx];
   d->kind = DkLnk;
   d->MshTyp = MshTyp;
   d->LnkTyp = LnkTyp;
   d->ItmTyp = typ;
   d->NmbItm = cnt;
   d->ItmSiz = OclTypSiz[typ];
   d->ItmLen = vec;
   d->NmbLin = g->NmbEle[MshTyp];
   d->LinSiz = (size_t)cnt * d->ItmSiz;
   d->MemSiz = (size_t)d->NmbLin * d->LinSiz;
   d->nam = nam ? nam : "";
   if(!AllocData(g, d)) return 0;
   g->LnkMat[MshTyp][LnkTyp] = idx;
   DbgPrint(g, "Allocated link data: index=%2d, size=%zu bytes\n", idx, d->MemSiz);
   return idx;
}

int GmlFreeData(size_t GmlIdx, int idx)
{
   GML_CTX(g, GmlIdx);
   if(idx < 1 || idx > GmlMaxDat || !g->dat[idx].use || !g->dat[idx].dev)
      return 0;
   // a slot of an uplink takes the whole uplink (its other slots and plans)
   for(int s = 0; s < GmlMaxEleTyp; s++)
      for(int t = 0; t < GmlMaxEleTyp; t++)
      {
         const Uplink &u = g->up[s][t];
         if(u.BalIdx && (idx == u.DegIdx || idx == u.BalIdx || idx == u.OffIdx))
         {
            DropUplink(g, s, t);
            return 1;
         }
      }
   Data *d = &g->dat[idx];
   if(d->kind == DkEle)
   {
      // Dropping the current table of a mesh kind also drops its ref table
      // (which the caller has no index for); links built on it are kept
      int t = d->MshTyp;
      if(g->TypIdx[t] == idx)
      {
         g->TypIdx[t] = 0;
         g->NmbEle[t] = 0;
         g->LnkMat[t][GmlVertices] = 0;
         if(g->RefIdx[t] && g->RefIdx[t] != idx) FreeDataSlot(g, g->RefIdx[t]);
         g->RefIdx[t] = 0;
      }
   }
   if(d->RedIdx) FreeDataSlot(g, d->RedIdx);
   if(d->OffIdx) FreeDataSlot(g, d->OffIdx);
   FreeDataSlot(g, idx);
   return 1;
}

// Line set: the host mirror is written; the whole item is uploaded when its
// last line is set (gmlib3.c:1894-1900).
int GmlSetDataLine(size_t GmlIdx, int idx, int lin, ...)
{
   GML_CTX(g, GmlIdx);
   if(idx < 1 || idx > GmlMaxDat || !g->dat[idx].use) return 0;
   Data *d = &g->dat[idx];
   if(lin < 0 || lin >= d->NmbLin || !EnsureHost(d)) return 0;
   int RefIdx = 0;
   va_list ap;
   va_start(ap, lin);

   if(d->kind == DkRaw || d->kind == DkArg)
   {
      void *p = va_arg(ap, void *);
      memcpy((char *)d->host + (size_t)lin * d->LinSiz, p, d->LinSiz);
   }
   else if(d->kind == DkLnk)
   {
      int w = d->NmbItm * d->ItmLen;
      int *tab = (int *)d->host;
      for(int i = 0; i < w; i++)
         tab[(size_t)lin * w + i] = va_arg(ap, int);
   }
   else if(d->kind == DkEle && d->MshTyp == GmlVertices)
   {
      RefIdx = g->RefIdx[GmlVertices];
      Data *r = &g->dat[RefIdx];
      if(!EnsureHost(r)) { va_end(ap); return 0; }
      float *c = (float *)d->host + (size_t)lin * 4;
      for(int i = 0; i < 3; i++)
         c[i] = (float)va_arg(ap, double);
      c[3] = 0.f;
      ((int *)r->host)[lin] = va_arg(ap, int);
   }
   else if(d->kind == DkEle)
   {
      RefIdx = g->RefIdx[d->MshTyp];
      Data *r = &g->dat[RefIdx];
      if(!EnsureHost(r)) { va_end(ap); return 0; }
      int siz = TypVecSiz[MshItmTyp[d->MshTyp]];
      int *e = (int *)d->host + (size_t)lin * siz;
      for(int i = 0; i < EleNmbNod[d->MshTyp]; i++)
         e[i] = va_arg(ap, int);
      ((int *)r->host)[lin] = va_arg(ap, int);
   }
   va_end(ap);

   if(lin == d->NmbLin - 1)
   {
      Upload(g, idx);
      if(RefIdx) Upload(g, RefIdx);
   }
   return 1;
}

// Line get: line 0 downloads the whole item (gmlib3.c:1921-1922); the other
// lines read the mirror.  Vertex lines are indexed by line (the reference's
// byte-offset indexing, gmlib3.c:1937, is not reproduced).
int GmlGetDataLine(size_t GmlIdx, int idx, int lin, ...)
{
   GML_CTX(g, GmlIdx);
   if(idx < 1 || idx > GmlMaxDat || !g->dat[idx].use) return 0;
   Data *d = &g->dat[idx];
   if(lin < 0 || lin >= d->NmbLin) return 0;
   if(lin == 0 || !d->host_valid)
   {
      Download(g, idx);
      if(d->kind == DkEle && g->RefIdx[d->MshTyp])
         Download(g, g->RefIdx[d->MshTyp]);
   }
   if(!d->host) return 0;
   va_list ap;
   va_start(ap, lin);

   if(d->kind == DkRaw || d->kind == DkArg || d->kind == DkRef)
   {
      void *p = va_arg(ap, void *);
      memcpy(p, (char *)d->host + (size_t)lin * d->LinSiz, d->LinSiz);
   }
   else if(d->kind == DkLnk)
   {
      int w = d->NmbItm * d->ItmLen;
      for(int i = 0; i < w; i++)
         *va_arg(ap, int *) = ((int *)d->host)[(size_t)lin * w + i];
   }
   else if(d->kind == DkEle && d->MshTyp == GmlVertices)
   {
      const float *c = (const float *)d->host + (size_t)lin * 4;
      for(int i = 0; i < 3; i++)
         *va_arg(ap, double *) = (double)c[i];
   }
   else if(d->kind == DkEle)
   {
      int siz = TypVecSiz[MshItmTyp[d->MshTyp]];
      const int *e = (const int *)d->host + (size_t)lin * siz;
      for(int i = 0; i < EleNmbNod[d->MshTyp]; i++)
         *va_arg(ap, int *) = e[i];
      Data *r = &g->dat[g->RefIdx[d->MshTyp]];
      int *ref = va_arg(ap, int *);
      if(ref && r->host) *ref = ((int *)r->host)[lin];
   }
   va_end(ap);
   return 1;
}

// Block set of mesh lines BegIdx..EndIdx (inclusive, as gmlib3.c:1990).  The
// user strides are inferred from the pointers to the first and last lines
// (gmlib3.c:1995-1996); vertex coordinates are floats.
int GmlSetDataBlock(size_t GmlIdx, int TypIdx, int BegIdx, int EndIdx,
                    void *DatBeg, void *DatEnd, int *RefBeg, int *RefEnd)
{
   GML_CTX(g, GmlIdx);
   if(TypIdx < 0 || TypIdx >= GmlMaxEleTyp) return 0;
   int DatIdx = g->TypIdx[TypIdx], RefIdx = g->RefIdx[TypIdx];
   if(!DatIdx) return 0;
   Data *d = &g->dat[DatIdx], *r = &g->dat[RefIdx];
   if(EndIdx <= BegIdx || d->kind != DkEle || BegIdx < 0 || EndIdx >= d->NmbLin) return 0;
   if(!EnsureHost(d) || !EnsureHost(r)) return 0;
   long long n = EndIdx - BegIdx;
   int *rt = (int *)r->host;

   if(d->MshTyp == GmlVertices)
   {
      const float *uc = (const float *)DatBeg;
      long long DatLen = ((const float *)DatEnd - (const float *)DatBeg) / n;
      long long RefLen = RefBeg ? (RefEnd - RefBeg) / n : 0;
      float *c = (float *)d->host;
      for(long long i = BegIdx; i <= EndIdx; i++)
      {
         for(int j = 0; j < 3; j++)
            c[i * 4 + j] = uc[(i - BegIdx) * DatLen + j];
         c[i * 4 + 3] = 0.f;
         if(RefBeg) rt[i] = RefBeg[(i - BegIdx) * RefLen];
      }
   }
   else
   {
      const int *ue = (const int *)DatBeg;
      long long DatLen = ((const int *)DatEnd - (const int *)DatBeg) / n;
      long long RefLen = RefBeg ? (RefEnd - RefBeg) / n : 0;
      int siz = TypVecSiz[MshItmTyp[d->MshTyp]], nn = EleNmbNod[d->MshTyp];
      int *e = (int *)d->host;
      for(long long i = BegIdx; i <= EndIdx; i++)
      {
         for(int j = 0; j < nn; j++)
            e[i * siz + j] = ue[(i - BegIdx) * DatLen + j];
         if(RefBeg) rt[i] = RefBeg[(i - BegIdx) * RefLen];
      }
   }
   if(EndIdx == d->NmbLin - 1)
   {
      Upload(g, DatIdx);
      Upload(g, RefIdx);
   }
   return 1;
}

// Block get (gmlib3.c:2042-2105).  Vertex lines return 3 floats per line
// (the reference writes 4 into a 3-float user stride, gmlib3.c:2076-2077).
int GmlGetDataBlock(size_t GmlIdx, int TypIdx, int BegIdx, int EndIdx,
                    void *DatBeg, void *DatEnd, int *RefBeg, int *RefEnd)
{
   GML_CTX(g, GmlIdx);
   if(TypIdx < 0 || TypIdx >= GmlMaxEleTyp) return 0;
   int DatIdx = g->TypIdx[TypIdx], RefIdx = g->RefIdx[TypIdx];
   if(!DatIdx) return 0;
   Data *d = &g->dat[DatIdx], *r = &g->dat[RefIdx];
   if(EndIdx <= BegIdx || d->kind != DkEle || BegIdx < 0 || EndIdx >= d->NmbLin) return 0;
   if(BegIdx == 0 || !d->host_valid || !r->host_valid)
   {
      Download(g, DatIdx);
      Download(g, RefIdx);
   }
   if(!d->host || !r->host) return 0;
   long long n = EndIdx - BegIdx;
   const int *rt = (const int *)r->host;

   if(d->MshTyp == GmlVertices)
   {
      float *uc = (float *)DatBeg;
      long long DatLen = ((float *)DatEnd - (float *)DatBeg) / n;
      long long RefLen = RefBeg ? (RefEnd - RefBeg) / n : 0;
      const float *c = (const float *)d->host;
      for(long long i = BegIdx; i <= EndIdx; i++)
      {
         // x, y, z only: a wider user stride keeps its other slots (reference Q21 wrote 4)
         for(int j = 0; j < 3; j++)
            uc[(i - BegIdx) * DatLen + j] = c[i * 4 + j];
         if(RefBeg) RefBeg[(i - BegIdx) * RefLen] = rt[i];
      }
   }
   else
   {
      int *ue = (int *)DatBeg;
      long long DatLen = ((int *)DatEnd - (int *)DatBeg) / n;
      long long RefLen = RefBeg ? (RefEnd - RefBeg) / n : 0;
      int siz = TypVecSiz[MshItmTyp[d->MshTyp]], nn = EleNmbNod[d->MshTyp];
      const int *e = (const int *)d->host;
      for(long long i = BegIdx; i <= EndIdx; i++)
      {
         for(int j = 0; j < nn; j++)
            ue[(i - BegIdx) * DatLen + j] = e[i * siz + j];
         if(RefBeg) RefBeg[(i - BegIdx) * RefLen] = rt[i];
      }
   }
   return 1;
}

int GmlUploadParameters(size_t GmlIdx)
{
   GML_CTX(g, GmlIdx);
   return g->ParIdx && Upload(g, g->ParIdx) ? 1 : 0;
}

int GmlDownloadParameters(size_t GmlIdx)
{
   GML_CTX(g, GmlIdx);
   return g->ParIdx && Download(g, g->ParIdx) ? 1 : 0;
}

void GmlSetCompilerOptions(size_t GmlIdx, char *OptStr)
{
   GML_CTX(g, GmlIdx);
   g->cflags = OptStr ? OptStr : "";
}

size_t GmlGetMemoryUsage(size_t GmlIdx)    { GML_CTX(g, GmlIdx); return g->MemSiz; }
size_t GmlGetMemoryTransfer(size_t GmlIdx) { GML_CTX(g, GmlIdx); return g->MovSiz; }
float  GmlGetMemoryAccess(size_t GmlIdx)   { GML_CTX(g, GmlIdx); return (float)g->MemAcc; }
float  GmlGetFlops(size_t GmlIdx)          { GML_CTX(g, GmlIdx); return (float)g->FltOpp; }
void   GmlDebugOn(size_t GmlIdx)           { GML_CTX(g, GmlIdx); g->DbgFlg = 1; }
void   GmlDebugOff(size_t GmlIdx)          { GML_CTX(g, GmlIdx); g->DbgFlg = 0; }
int    GmlCheckFP64(size_t GmlIdx)         { GML_CTX(g, GmlIdx); return g->DblExt; }

void GmlIncludeUserToolkit(size_t GmlIdx, char *src)
{
   GML_CTX(g, GmlIdx);
   g->UsrTlk = src;
}

int GmlGetMeshInfo(size_t GmlIdx, int typ, int *NmbLin, int *DatIdx)
{
   GML_CTX(g, GmlIdx);
   if(typ < 0 || typ >= GmlMaxEleTyp) return 0;
   if(NmbLin) *NmbLin = 0;
   if(DatIdx) *DatIdx = 0;
   int idx = g->TypIdx[typ];
   if(!idx || !g->dat[idx].NmbLin) return 0;
   if(NmbLin) *NmbLin = (int)g->dat[idx].NmbLin;
   if(DatIdx) *DatIdx = idx;
   return 1;
}

// Link lengths and widths: n x w base rows, N x W high-degree rows.
int GmlGetLinkInfo(size_t GmlIdx, int SrcTyp, int DstTyp, int *n, int *w, int *N, int *W)
{
   GML_CTX(g, GmlIdx);
   if(SrcTyp < 0 || SrcTyp >= GmlMaxEleTyp || DstTyp < 0 || DstTyp >= GmlMaxEleTyp) return 0;
   Uplink *u = &g->up[SrcTyp][DstTyp];
   if(u->BalIdx)
   {
      *n = u->MaxPos;
      *w = u->BasWid;
      *N = u->HghWid ? (int)(g->NmbEle[SrcTyp] - u->MaxPos) : 0;
      *W = u->HghWid;
      return 1;
   }
   int idx = g->LnkMat[SrcTyp][DstTyp];
   if(!idx) return 0;
   *n = (int)g->dat[idx].NmbLin;
   *w = g->dat[idx].NmbItm * g->dat[idx].ItmLen;
   *N = *W = 0;
   return 1;
}

double GmlGetWallClock()
{
   struct timeval tp;
   gettimeofday(&tp, nullptr);
   return tp.tv_sec + tp.tv_usec / 1000000.;
}

// Cache-reuse estimate of the element numbering (gmlib3.c:4409-4442): a
// 16-entry direct-mapped tag table over 1024-index blocks probed with the
// first node of every element of every mesh kind.
float GmlEvaluateNumbering(size_t GmlIdx)
{
   GML_CTX(g, GmlIdx);
   int tag[64] = {0};
   long long hit = 0, mis = 0;
   for(int typ = GmlEdges; typ < GmlMaxEleTyp; typ++)
   {
      int idx = g->TypIdx[typ];
      if(!idx) continue;
      Data *d = &g->dat[idx];
      if(!d->host_valid) Download(g, idx);
      const int *e = (const int *)d->host;
      if(!e) continue;
      for(long long i = 0; i < d->NmbLin; i++)
      {
         int v = e[i * d->ItmLen];
         int pos = (v >> 4) & 0xf, adr = v >> 10;
         if(tag[pos] == adr) hit++;
         else { mis++; tag[pos] = adr; }
      }
   }
   if(hit + mis == 0) return 0.f;
   return 100.f * (float)hit / (float)(hit + mis);
}

void *GmlxGetDevicePointer(size_t GmlIdx, int DatIdx, size_t *bytes)
{
   GML_CTX(g, GmlIdx);
   if(DatIdx < 1 || DatIdx > GmlMaxDat || !g->dat[DatIdx].use) return nullptr;
   if(bytes) *bytes = g->dat[DatIdx].MemSiz;
   return g->dat[DatIdx].dev;
}

int GmlxSync(size_t GmlIdx)
{
   GML_CTX(g, GmlIdx);
   if(g->host_only) return 0;
   return hipStreamSynchronize(g->stream) == hipSuccess ? 1 : 0;
}

void *GmlxGetStream(size_t GmlIdx)
{
   GML_CTX(g, GmlIdx);
   return (void *)g->stream;
}

int GmlxInvalidateHost(size_t GmlIdx, int DatIdx)
{
   GML_CTX(g, GmlIdx);
   if(DatIdx < 1 || DatIdx > GmlMaxDat || !g->dat[DatIdx].use) return 0;
   g->dat[DatIdx].host_valid = false;
   BumpVersion(&g->dat[DatIdx]);
   return 1;
}

} // extern "C"
