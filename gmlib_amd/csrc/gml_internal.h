// Internal structures of the MI355X-native GMlib engine (libgml.so).
//
// One GmlCtx per GmlInit handle.  Every data item lives in HBM for the life
// of the handle; the host mirror is allocated on first line/block access.
// All device work of a handle is issued on its own HIP stream in order.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>
#include <mutex>
#include <array>

#include "../../include/gmlib3.h"
#include "../../include/gmlx.h"

namespace gml {

// Data kinds (the reference's data_type enum, gmlib3.c:71-72)
enum DatKind { DkArg = 0, DkRaw, DkLnk, DkEle, DkRef, DkMat, DkVec, DkBal, DkOff };

// Tables whose values are part of the interface (gmlib3.c:170-365):
// sizes of the OpenCL types, element node counts, link widths.
extern const int OclTypSiz[GmlMaxOclTyp];
extern const int TypVecSiz[GmlMaxOclTyp];
extern const char *OclTypStr[GmlMaxOclTyp];
extern const char *OclNulVec[GmlMaxOclTyp];
extern const int MshItmTyp[GmlMaxEleTyp];
extern const int EleNmbNod[GmlMaxEleTyp];
extern const int MshTypDim[GmlMaxEleTyp];
extern const char *BalTypStr[GmlMaxEleTyp];
extern const char *MshTypStr[GmlMaxEleTyp];
extern const char *MshRefStr[GmlMaxEleTyp];
extern const int LenMatBas[GmlMaxEleTyp][GmlMaxEleTyp];
extern const int LenMatMax[GmlMaxEleTyp][GmlMaxEleTyp];
extern const int NmbTpoLnk[GmlMaxEleTyp][GmlMaxEleTyp];
extern const int ItmEdgNod[8][12][2];
extern const int ItmFacDeg[8][6];
extern const int ItmFacNod[8][6][4];

struct Data {
   bool     use = false;
   int      kind = 0, MshTyp = 0, LnkTyp = 0, ItmTyp = 0;
   int      NmbItm = 0, ItmSiz = 0, ItmLen = 0;
   long long NmbLin = 0;
   size_t   LinSiz = 0, MemSiz = 0;
   void    *dev = nullptr;
   void    *host = nullptr;
   std::string nam, VoyNam;
   const char *src = nullptr;        // parameter struct source (caller-owned, kept by pointer)
   int      RedIdx = 0;
   int      OffIdx = 0;              // DkBal: slot of the chunk offset table
   bool     host_valid = false;      // mirror holds the device contents
   unsigned long long ver = 0;       // changes whenever the device contents may have changed
};

// give a data item a new, process-wide unique version (its device contents
// changed: upload, kernel write, allocation, caller invalidation)
void BumpVersion(Data *d);

// Vertex/entity -> element uplink in SELL-64 layout: rows are grouped in
// chunks of 64 consecutive source entities; chunk c stores its entries
// slab-major at ent[(off[c] + k) * 64 + lane], k < width of the chunk.
// Tile plan of an LDS-staged uplink gather in generated kernels: for tile t
// (rows t*grp .. t*grp+grp-1 of the source kind), the distinct destination
// entities of its ball rows, ascending, and for every ball slot (SELL-64
// layout of the ball) the tile-local index of its entity.
struct StagePlan {
   int grp = 0, max_nt = 0;
   long long nrows = 0, ntile = 0;
   unsigned int *tofs = nullptr;    // [ntile+1]
   int *tlist = nullptr;            // destination ids, tile by tile
   unsigned short *lidx = nullptr;  // per ball slot
   size_t bytes = 0;
};

struct Uplink {
   int   DegIdx = 0, BalIdx = 0, OffIdx = 0;
   int   BasWid = 0, HghWid = 0, MaxPos = 0, MaxDeg = 0;
   long long nslab = 0;
   long long SrcCnt = 0;                   // source entities it was built for
   unsigned long long DstVer = 0;          // version of the element table it was built from
   unsigned long long SrcVer = 0;          // version of the source table (0 for vertices)
   bool  pinned = false;                   // held by a compiled kernel or handed to the caller
   StagePlan *stage[3] = {nullptr, nullptr, nullptr};   // work-group size 64, 128, 256
};

// a pair of timing events around one launch, from the process's event pool
// of the device they were created on (NewRec / FreeRec)
struct LaunchRec { hipEvent_t beg = nullptr, end = nullptr; int dev = 0; };

struct Kernel {
   bool          use = false;
   std::string   name, src;
   hipModule_t   mod = nullptr;
   hipFunction_t fn = nullptr;
   std::vector<int> args;            // data slot per pointer argument (DkBal expands to 2)
   std::vector<int> wargs;           // slots the kernel writes
   long long     NmbLin[2] = {0, 0}; // {count, offset}
   int           HghIdx = 0;
   bool          has_par = true;
   // launch shape autotuning (group size x XCD remap x form); forms: 0 one
   // row per thread, 1 paired (two rows), 2 element tiles (LDS-staged
   // element -> vertex data, kDownTile rows per work-group)
   int           grp = 256, remap = 1, tuned = 0, tune_step = 0, form = 0;
   std::vector<std::array<int, 3>> shapes;   // the tuner's candidates {group, remap, form}
   // the body names a work-item builtin (get_global_id ...): its launches
   // keep the plain block order, so get_global_id(0) is the row
   bool          rowids = false;
   // paired form (two rows per thread), when the kernel has one
   hipModule_t   mod2 = nullptr;
   hipFunction_t fn2 = nullptr;
   std::string   src2;
   // element-tile form, when the kernel has one: the element kind whose
   // table is staged and the byte size of each staged item
   hipModule_t   mod3 = nullptr;
   hipFunction_t fn3 = nullptr;
   std::string   src3;
   int           dtyp = -1;
   std::vector<int> dsiz;
   // LDS-staged uplink data (stage_src >= 0): the uplink, the byte size of
   // each staged item, the work-group size its plan was built for
   int           stage_src = -1, stage_dst = -1, stage_grp = 0;
   std::vector<int> stage_siz;
   std::vector<LaunchRec> tune_ev;
   // profiling
   std::vector<LaunchRec> evts;
   double        run_time = 0.;
};

struct Matrix;
struct Batch;
struct Vector;

// Tile plan of the fused smoothing sweep (gml_plan.cpp)
struct SweepPlan {
   int TV = 0;
   long long ntile = 0;
   int max_nv = 0, max_nt = 0;              // largest per-tile counts (LDS sizing)
   long long sum_nv = 0, sum_nt = 0;        // totals over tiles (halo redundancy)
   long long sum_ring = 0;                  // operator form: total 1-ring slots (V + 2E)
   unsigned int *vofs = nullptr;            // [ntile+1] offsets into vlist
   int *vlist = nullptr;                    // distinct vertices per tile, ascending
   unsigned int *tofs = nullptr;            // [ntile+1] offsets into tloc
   unsigned short *tloc = nullptr;          // 4 tile-local vertex indices per distinct tet
   unsigned short *lidx = nullptr;          // ball slot -> tile-local tet index (SELL-64)
   // operator form: per vertex, its 1-ring without itself as 16-bit
   // (tile-local index << 4 | count of ball tets containing the neighbour),
   // SELL-64 over the vertices (deg v = sum of the ring's counts / 3)
   unsigned short *nbr = nullptr;
   unsigned int *noff = nullptr;
   long long nslab = 0;
   size_t bytes = 0;
};

struct ScratchBlock { void *p; size_t bytes; bool busy; };

// Tile plan of an element -> vertex gather staged in LDS (gml_plan.cpp):
// tiles of `rows` consecutive elements (the largest of 512, 256, 128, 64
// whose tiles each reach at most 256 distinct vertices: 512 for a
// Hilbert-ordered tet mesh); per tile the ascending distinct vertex ids of
// its rows (every slot of the element table's rows, as the generated reads
// take them); per row its slots' tile-local indices, one byte each, packed
// four to a word.  No plan when even 64-row tiles exceed 256 vertices or an
// id is outside the vertex table.  Valid while the element table keeps the
// version it was built from.
const int kDownTile = 512;                  // largest tile
struct DownPlan {
   int typ = 0, w = 0, wq = 0, rows = 0;    // element kind, slots per row, index words per row, tile rows
   long long nrows = 0, ntile = 0;
   int max_nv = 0;                          // largest tile list
   unsigned long long ver = 0;              // version of the element table
   const void *src = nullptr;               // its device address
   unsigned int *tofs = nullptr;            // [ntile + 1]
   int *tlist = nullptr;                    // vertex ids, tile by tile
   unsigned int *lid = nullptr;             // [nrows * wq]
   long long sum_nv = 0;
   size_t bytes = 0;
   bool none = false;                       // not stageable (a tile over 256 vertices, or bad ids)
};

// launch shape of a built-in kernel (work-group size, XCD block remap,
// items per thread where a kernel takes several), per handle
const int kNmbBuiltin = 16;
struct Shape { int grp = 256, remap = 1, per = 2; };

struct GmlCtx {
   int          dev = 0;
   hipStream_t  stream = nullptr;
   int          NmbKrn = 0, ParIdx = 0, DbgFlg = 0, DblExt = 1;
   int          TypIdx[GmlMaxEleTyp] = {0};
   int          RefIdx[GmlMaxEleTyp] = {0};
   long long    NmbEle[GmlMaxEleTyp] = {0};
   int          LnkMat[GmlMaxEleTyp][GmlMaxEleTyp] = {{0}};
   Uplink       up[GmlMaxEleTyp][GmlMaxEleTyp];
   int          RedKrn[GmlMaxRed] = {0};
   const char  *UsrTlk = nullptr;
   std::string  cflags;
   size_t       MemSiz = 0, MovSiz = 0;
   double       MemAcc = 0., FltOpp = 0.;
   Data         dat[GmlMaxDat + 1];
   Kernel       krn[GmlMaxKrn + 1];
   Matrix      *mat[GmlMaxMat + 1] = {nullptr};
   Vector      *vec[GmlMaxVec + 1] = {nullptr};
   // reduction scratch (device): partials, ticket, result
   void        *red_scratch = nullptr;
   double      *red_host = nullptr;  // pinned result slot
   std::vector<LaunchRec> red_evts[GmlMaxRed];
   double       red_time[GmlMaxRed] = {0.};
   void        *dummy_par = nullptr; // parameter buffer when none was declared
   // device budget (0 = unlimited): used by the serving admission path
   size_t       budget = 0;
   // fused smoothing: tile plan and the second coordinate buffer
   SweepPlan   *plan = nullptr;             // exact fused form
   SweepPlan   *oplan = nullptr;            // operator form
   void        *xalt = nullptr;
   long long    xalt_n = 0;                 // vertices xalt was allocated for
   // batches of meshes processed together (gml_batch.hip)
   Batch       *bat[GmlxMaxBatch + 1] = {nullptr};
   // host-only handle (GmlxInitHostOnly): data in host memory, kernels are
   // generated and compiled but never launched
   bool         host_only = false;
   // cached temporaries of the item kernels (gml_items.hip)
   std::vector<ScratchBlock> iscratch;
   // tuning state of this handle: built-in launch shapes
   // (GmlxSetBuiltinShape / Items), sweep form (GmlxSetSmoothMode), LDS
   // staging of uplink data in kernels compiled later (GmlxSetStaging)
   Shape        shape[kNmbBuiltin];
   int          smooth_mode = 2;
   int          stage_uplinks = 1;
   // element -> vertex tile plans, per element kind (built on first use)
   DownPlan    *dplan[GmlMaxEleTyp] = {nullptr};
   unsigned long long dplan_seen[GmlMaxEleTyp] = {0};   // table version of the last unplanned request
   // the hand-written quality kernel defaults to its tile-staged form
   // (items per thread 0) in work-groups of 128
   GmlCtx() { shape[GmlxBuiltinQuality] = Shape{128, 1, 0}; }
};

// --- helpers ---------------------------------------------------------------
#define GML_CTX(g, h) gml::GmlCtx *g = reinterpret_cast<gml::GmlCtx *>(h)
#define HIPCHK(x) do { hipError_t e_ = (x); if(e_ != hipSuccess) { \
   fprintf(stderr, "gml: HIP error %s at %s:%d (%s)\n", hipGetErrorString(e_), __FILE__, __LINE__, #x); } } while(0)

int  NewDatIdx(GmlCtx *g);
bool AllocData(GmlCtx *g, Data *d, bool zero = true);
int NewMeshTable(GmlCtx *g, int MshTyp, int NmbLin, bool zero);
void FreeDataSlot(GmlCtx *g, int idx);
bool EnsureHost(Data *d);
size_t Upload(GmlCtx *g, int idx);
size_t Download(GmlCtx *g, int idx);
int  GetCntVec(int siz, int *cnt, int *vec, int *typ);
void *DevAlloc(GmlCtx *g, size_t bytes);
void DevFree(GmlCtx *g, void *p);
bool DevCopy(GmlCtx *g, void *dst, const void *src, size_t bytes, int dir);  // dir 1: H->D, 2: D->H, 3: D->D
void DbgPrint(GmlCtx *g, const char *fmt, ...);

// topology (gml_topology.cpp / gml_topology_kernels.hip)
int  BuildUplink(GmlCtx *g, int SrcTyp, int DstTyp);
// the uplink of the current tables: built if absent; one built from other
// table contents or another source count is dropped and built again, unless
// it is pinned (then null)
Uplink *CurrentUplink(GmlCtx *g, int SrcTyp, int DstTyp);
// free an uplink's data slots and the plans built on it, and forget it
void DropUplink(GmlCtx *g, int SrcTyp, int DstTyp);
int  BuildDownlink(GmlCtx *g, int SrcTyp, int DstTyp);
int  gpu_vertex_ball(hipStream_t s, const int *ele, int stride, int nnod, long long nele,
                     long long nver, int *deg, unsigned int **off_out, int **ent_out,
                     long long *nslab_out, int *maxdeg_out, int *firsthigh_out, int baswid);

// SELL-64 layout of rows of the given device degrees (gml_topology.hip)
int  gpu_sell_layout(hipStream_t s, const int *deg, long long n, int baswid, unsigned int **off_out,
                     int **ent_out, long long *nslab_out, int *maxdeg_out, int *firsthigh_out);

// topology from element items on the device (gml_items.hip): extraction of
// the distinct items of degree deg into the table of kind typ (count, -1 on
// failure); downlink table Src -> Dst (data index, 0 on failure); shell rows
// of Src entities over Dst elements; neighbours of kind typ across items of
// kind ngb into link data idx of rw ints per line
int  GpuExtractItems(size_t GmlIdx, GmlCtx *g, int typ, int deg);
int  GpuDownlink(GmlCtx *g, int Src, int Dst, const std::string &nam);
int  GpuShell(GmlCtx *g, int Src, int Dst, int *degdev, unsigned int **off, int **ent, long long *nslab,
              int *maxdeg, int *firsthigh);
int  GpuNeighbours(GmlCtx *g, int typ, int ngb, int idx, int rw);
bool ItemAlloc(GmlCtx *g, void **p, size_t bytes);
void ItemFree(GmlCtx *g, void *p);
void FreeItemPool(GmlCtx *g, bool idle_only = false);

// reductions (gml_reduce.hip)
int  ReduceDevice(GmlCtx *g, const float *x, long long n, int op, double *res, bool timed);
size_t ReduceScratchBytes();
int  NoAllocScratchInit(int dev);   // device current; once per device (GmlInit)

// block algebra (gml_algebra.hip)
void FreeAlgebra(GmlCtx *g);

void HarvestEvents(std::vector<LaunchRec> &ev, double *acc);
// take an event pair from the pool of the current device (created when the
// pool is empty); false if none could be created
bool NewRec(LaunchRec *r);
// return an event pair to its device's pool
void FreeRec(const LaunchRec &r);

// launch shape of a built-in kernel on a handle
inline Shape BuiltinShape(const GmlCtx *g, int which) { return g->shape[which]; }

// Block index of this work-group in launch order, or, with remap, remapped
// so that each of the 8 XCDs (which take dispatched blocks round-robin) runs
// one contiguous range of blocks and its L2 sees neighbouring data.
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

// built-in kernel timing: HIP event pairs per launch, per handle and kernel
struct KernStat {
   std::vector<LaunchRec> ev;
   double time = 0.;
   long long launches = 0;
};
KernStat *Stat(GmlCtx *g, int which);

// Record an event pair around launch() on the handle's stream.
template<class F>
int TimedLaunch(GmlCtx *g, int which, F launch)
{
   KernStat *st = Stat(g, which);
   LaunchRec r;
   bool ev = NewRec(&r);
   if(ev) (void)hipEventRecord(r.beg, g->stream);
   launch();
   if(hipGetLastError() != hipSuccess)
   {
      if(ev) FreeRec(r);
      return -6;
   }
   if(ev)
   {
      (void)hipEventRecord(r.end, g->stream);
      st->ev.push_back(r);
      if(st->ev.size() > 256) HarvestEvents(st->ev, &st->time);
   }
   st->launches++;
   return 1;
}

// built-in mesh kernels (gml_mesh_kernels.hip, gml_plan.cpp, gml_batch.hip)
void FreeBuiltinStats(GmlCtx *g);
void FreeBatches(GmlCtx *g);
SweepPlan *NewSweepPlan(GmlCtx *g, int TV, int op);
void FreeSweepPlan(GmlCtx *g, SweepPlan *p);
StagePlan *NewStagePlan(GmlCtx *g, int Src, int Dst, int grp, long long nrows);
void FreeStagePlan(GmlCtx *g, StagePlan *p);
// GPU builders of the element-tile and staging plans (gml_plan_gpu.hip)
int GpuTilePlan(GmlCtx *g, const int *ele, long long n, int w, int wq, long long nv, int TR, unsigned int **tofs,
                int **tlist, unsigned int **lid, long long *sum, int *maxn);
int GpuStagePlan(GmlCtx *g, const int *ent, const unsigned int *off, const int *deg, long long nrows, int grp,
                 size_t nslot, unsigned int **tofs, int **tlist, unsigned short *lidx, long long *sum, int *maxn);
int GpuSweepOpPlan(GmlCtx *g, const int *ent, const unsigned int *off, const int *deg, const int *tet, long long nv,
                   int TV, unsigned int **vofs, int **vlist, unsigned int **noff, unsigned short **nbr, long long *nslab,
                   long long *sum_halo, int *max_halo, long long *sum_ring);
// the current element -> vertex tile plan of element kind typ, rebuilt when
// the element table changed; nullptr when the kind cannot be staged
DownPlan *GetDownPlan(GmlCtx *g, int typ);
// the plan if it exists, or built when the same table version was asked for
// before (a table used once never pays for a plan); nullptr otherwise
DownPlan *RepeatDownPlan(GmlCtx *g, int typ);
void FreeDownPlan(GmlCtx *g, DownPlan *p);

// launcher (gml_codegen.cpp)
int  RunKernel(GmlCtx *g, Kernel *k);
void SetShapes(Kernel *k);   // the launch tuner's candidates of a compiled kernel

} // namespace gml
