// GmlReduceVector: single-pass wave64 reductions over a float vector.
//
// Operation meanings follow the reference (sources/gmlib3.c:3229-3332,
// sources/reduce.cl:6-137): min, max, sum, L0 = number of nonzero entries,
// L1 = sum |x|, L2 = sum x^2 (no square root), Linf = max |x|.  Unlike the
// reference, which leaves one partial per work-group for a float loop on the
// host and drops the last partial group, the whole vector is reduced on the
// GPU: each thread streams float4 loads, min/max stay exact in FP32 and the
// sums accumulate in FP64; waves combine with cross-lane shuffles, work-groups
// through LDS and store one partial each, and a one-work-group kernel
// combines the partials in block order (deterministic, and free of the
// per-work-group device-scope fence a last-block ticket would need, which on
// CDNA writes back and invalidates the XCD's L2).  Only the final double
// crosses PCIe.

#include "gml_internal.h"
#include "gml_device.h"

#include <mutex>

namespace gml {

static const int kRedBlock = 256;
static const int kRedMaxGrid = 16384;

struct RedScratch {
   double part[kRedMaxGrid];
   double result;
};

size_t ReduceScratchBytes() { return sizeof(RedScratch) + 256; }

template<int OP> __device__ __forceinline__ double red_init()
{
   return OP == GmlMin ? 3.402823466e+38 : (OP == GmlMax ? -3.402823466e+38 : 0.);
}

template<int OP> __device__ __forceinline__ double red_comb(double a, double b)
{
   if(OP == GmlMin) return fmin(a, b);
   if(OP == GmlMax || OP == GmlLinf) return fmax(a, b);
   return a + b;
}

template<int OP> __device__ __forceinline__ double red_elem(float x)
{
   if(OP == GmlL0) return x != 0.f ? 1. : 0.;
   if(OP == GmlL1) return fabs((double)x);
   if(OP == GmlL2) return (double)x * (double)x;
   if(OP == GmlLinf) return (double)fabsf(x);
   return (double)x;
}

template<int OP> __device__ __forceinline__ double wave_reduce(double v)
{
   for(int o = 32; o > 0; o >>= 1)
      v = red_comb<OP>(v, __shfl_xor(v, o, 64));
   return v;
}

template<int OP> __device__ __forceinline__ double block_reduce(double v, double *lds)
{
   v = wave_reduce<OP>(v);
   const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
   if(lane == 0) lds[wid] = v;
   __syncthreads();
   v = threadIdx.x < (unsigned)nw ? lds[threadIdx.x] : red_init<OP>();
   if(wid == 0) v = wave_reduce<OP>(v);
   return v;
}

// min/max accumulate in float registers (exact); sums in double.  A
// work-group reads one contiguous block of kRedLoads x 256 float4 per pass
// (thread t takes float4 t, t + 256, ... of the block), every load of the pass
// in flight; a slot past the end loads nothing and contributes the
// operation's identity, which leaves every result unchanged.  The grid holds
// one block per work-group (up to 16384 work-groups, which then stride over
// the rest), so a vector streams in one memory round trip.  Measured on
// 100.7M floats: 65 us per GmlReduceVector (6.2 TB/s, the finishing kernel
// included), from 74 us with grid-strided loads four at a time.
static const int kRedLoads = 8;

template<int OP> __device__ __forceinline__ float red_fold(float a, v4f v)
{
   if(OP == GmlMin) return fminf(a, fminf(fminf(v.x, v.y), fminf(v.z, v.w)));
   if(OP == GmlMax) return fmaxf(a, fmaxf(fmaxf(v.x, v.y), fmaxf(v.z, v.w)));
   return fmaxf(a, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))));
}

template<int OP>
__global__ void __launch_bounds__(kRedBlock) k_reduce(const float *__restrict__ x, long long n, RedScratch *__restrict__ s)
{
   __shared__ double lds[kRedBlock / 64];
   const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   const long long nth = (long long)gridDim.x * blockDim.x;
   double acc = red_init<OP>();
   const long long n4 = n >> 2;
   const v4f *x4 = reinterpret_cast<const v4f *>(x);

   if(OP == GmlMin || OP == GmlMax || OP == GmlLinf)
   {
      const float id = (float)red_init<OP>();
      float a = id;
      for(long long i = (long long)blockIdx.x * (kRedLoads * kRedBlock) + threadIdx.x; i < n4; i += kRedLoads * nth)
      {
         v4f v[kRedLoads];
#pragma unroll
         for(int u = 0; u < kRedLoads; u++)
            v[u] = i + u * kRedBlock < n4 ? __builtin_nontemporal_load(&x4[i + u * kRedBlock]) : (v4f){id, id, id, id};
#pragma unroll
         for(int u = 0; u < kRedLoads; u++) a = red_fold<OP>(a, v[u]);
      }
      for(long long i = n4 * 4 + tid; i < n; i += nth)
      {
         float v = x[i];
         if(OP == GmlMin)      a = fminf(a, v);
         else if(OP == GmlMax) a = fmaxf(a, v);
         else                  a = fmaxf(a, fabsf(v));
      }
      acc = (double)a;
   }
   else
   {
      double a0 = 0., a1 = 0.;
      for(long long i = (long long)blockIdx.x * (kRedLoads * kRedBlock) + threadIdx.x; i < n4; i += kRedLoads * nth)
      {
         v4f v[kRedLoads];
#pragma unroll
         for(int u = 0; u < kRedLoads; u++)
            v[u] = i + u * kRedBlock < n4 ? __builtin_nontemporal_load(&x4[i + u * kRedBlock]) : (v4f){0.f, 0.f, 0.f, 0.f};
#pragma unroll
         for(int u = 0; u < kRedLoads; u++)
         {
            a0 += red_elem<OP>(v[u].x) + red_elem<OP>(v[u].y);
            a1 += red_elem<OP>(v[u].z) + red_elem<OP>(v[u].w);
         }
      }
      for(long long i = n4 * 4 + tid; i < n; i += nth)
         a0 += red_elem<OP>(x[i]);
      acc = a0 + a1;
   }

   acc = block_reduce<OP>(acc, lds);
   if(threadIdx.x == 0) s->part[blockIdx.x] = acc;
}

// one work-group of kFinBlock threads: combine the np partials in a fixed
// order, eight independent loads in flight per thread
static const int kFinBlock = 1024;

// host, if not null, is pinned host memory that also receives the result (a
// system-scope store: visible to the host once the stream has synchronised),
// which saves the device-to-host copy of one double, 3.6 us per call
template<int OP>
__global__ void __launch_bounds__(kFinBlock) k_reduce_finish(RedScratch *__restrict__ s, int np, double *host)
{
   __shared__ double lds[kFinBlock / 64];
   double a = red_init<OP>();
   for(int base = 0; base < np; base += 8 * kFinBlock)
   {
      double v[8];
#pragma unroll
      for(int u = 0; u < 8; u++)
      {
         const int i = base + u * kFinBlock + (int)threadIdx.x;
         v[u] = i < np ? s->part[i] : red_init<OP>();
      }
#pragma unroll
      for(int u = 0; u < 8; u++) a = red_comb<OP>(a, v[u]);
   }
   a = block_reduce<OP>(a, lds);
   if(threadIdx.x == 0)
   {
      s->result = a;
      if(host) __hip_atomic_store(host, a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
   }
}

// work-groups of a reduction over n floats (one partial each)
static int RedGrid(long long n)
{
   const long long want = (n / 4 + kRedBlock * kRedLoads - 1) / (kRedBlock * kRedLoads);
   return (int)(want < 1 ? 1 : (want > kRedMaxGrid ? kRedMaxGrid : want));
}

static int ReduceLaunch(hipStream_t st, const float *x, long long n, int op, RedScratch *s, double *host = nullptr)
{
   // about four float4 per thread, up to 16384 work-groups
   const int grid = RedGrid(n);
#define GML_RED(O) do { hipLaunchKernelGGL(k_reduce<O>, dim3(grid), dim3(kRedBlock), 0, st, x, n, s); \
                         hipLaunchKernelGGL(k_reduce_finish<O>, dim3(1), dim3(kFinBlock), 0, st, s, grid, host); } while(0)
   switch(op)
   {
      case GmlMin:  GML_RED(GmlMin);  break;
      case GmlMax:  GML_RED(GmlMax);  break;
      case GmlSum:  GML_RED(GmlSum);  break;
      case GmlL0:   GML_RED(GmlL0);   break;
      case GmlL1:   GML_RED(GmlL1);   break;
      case GmlL2:   GML_RED(GmlL2);   break;
      case GmlLinf: GML_RED(GmlLinf); break;
      default: return -3;
   }
#undef GML_RED
   return hipGetLastError() == hipSuccess ? 1 : -6;
}

// ---------------------------------------------------------------------------
// Scratch of the no-allocation form (GmlxReduceVectorNoAlloc): the partials
// and the result live in this file-scope device array instead of the
// handle's scratch.  Every GPU the library runs on holds its own copy
// (sizeof(RedScratch) = 131,080 bytes) until the process ends.  In host code
// the name `scratch` is a host-side stand-in, not a device address, so it is
// never passed to a kernel: NoAllocScratchInit looks up each device's address
// once, with that device current, and kernels take that address.
// ---------------------------------------------------------------------------
static __device__ __align__(256) RedScratch scratch[1];

static const int kScrMaxDev = 64;
static std::mutex g_scr_mtx;
static RedScratch *g_scr_ptr[kScrMaxDev] = {nullptr};

int NoAllocScratchInit(int dev)
{
   if(dev < 0 || dev >= kScrMaxDev) return 0;
   std::lock_guard<std::mutex> lk(g_scr_mtx);
   if(g_scr_ptr[dev]) return 1;
   void *p = nullptr;
   const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
   if(e != hipSuccess || !p)
   {
      fprintf(stderr, "gml: no address for the reduction scratch on device %d: %s\n", dev, hipGetErrorString(e));
      (void)hipGetLastError();
      return 0;
   }
   g_scr_ptr[dev] = (RedScratch *)p;
   return 1;
}

static RedScratch *NoAllocScratch(int dev)
{
   if(dev < 0 || dev >= kScrMaxDev) return nullptr;
   std::lock_guard<std::mutex> lk(g_scr_mtx);
   return g_scr_ptr[dev];
}

// Reduce n floats at device address x; *res receives the result.
int ReduceDevice(GmlCtx *g, const float *x, long long n, int op, double *res, bool timed)
{
   if(g->host_only) return -6;
   RedScratch *s = (RedScratch *)g->red_scratch;
   LaunchRec r;
   bool ev = timed && NewRec(&r);
   if(ev) (void)hipEventRecord(r.beg, g->stream);
   int rc = ReduceLaunch(g->stream, x, n, op, s, g->red_host);
   if(rc != 1)
   {
      if(ev) FreeRec(r);
      return rc;
   }
   if(ev)
   {
      (void)hipEventRecord(r.end, g->stream);
      g->red_evts[op].push_back(r);
   }
   if(hipStreamSynchronize(g->stream) != hipSuccess)
      return -6;
   *res = *(volatile double *)g->red_host;
   if(g->red_evts[op].size() > 256) HarvestEvents(g->red_evts[op], &g->red_time[op]);
   return 1;
}

} // namespace gml

using namespace gml;

extern "C" {

int GmlReduceVector(size_t GmlIdx, int DatIdx, int RedOpp, double *nrm)
{
   GML_CTX(g, GmlIdx);
   if(DatIdx < 1 || DatIdx > GmlMaxDat || !g->dat[DatIdx].use)
   {
      printf("Invalid data index: %d\n", DatIdx);
      return -1;
   }
   Data *d = &g->dat[DatIdx];
   if(d->ItmTyp != GmlFlt || d->NmbItm != 1 || d->ItmLen != 1)
   {
      printf("Invalid data structure: count %d, type %s, length %d\n", d->NmbItm, OclTypStr[d->ItmTyp], d->ItmLen);
      return -2;
   }
   if(RedOpp < 0 || RedOpp >= GmlMaxRed)
   {
      printf("Invalid operation code %d\n", RedOpp);
      return -3;
   }
   if(!nrm) return -4;
   return ReduceDevice(g, (const float *)d->dev, d->NmbLin, RedOpp, nrm, true);
}

// GmlReduceVector without any allocation at call time: the same kernels
// and result, with the partials in the file-scope scratch of the handle's
// device; no timing events.  Calls that overlap on one GPU race on that
// scratch.  Every launch writes each partial its finish reads, so nothing
// has to be cleared between calls.
int GmlxReduceVectorNoAlloc(size_t GmlIdx, int DatIdx, int RedOpp, double *nrm)
{
   GML_CTX(g, GmlIdx);
   if(DatIdx < 1 || DatIdx > GmlMaxDat || !g->dat[DatIdx].use)
   {
      printf("Invalid data index: %d\n", DatIdx);
      return -1;
   }
   Data *d = &g->dat[DatIdx];
   if(d->ItmTyp != GmlFlt || d->NmbItm != 1 || d->ItmLen != 1)
   {
      printf("Invalid data structure: count %d, type %s, length %d\n", d->NmbItm, OclTypStr[d->ItmTyp], d->ItmLen);
      return -2;
   }
   if(RedOpp < 0 || RedOpp >= GmlMaxRed)
   {
      printf("Invalid operation code %d\n", RedOpp);
      return -3;
   }
   if(!nrm) return -4;
   if(g->host_only) return -6;
   RedScratch *s = NoAllocScratch(g->dev);
   if(!s)
   {
      printf("No reduction scratch on device %d\n", g->dev);
      return -6;
   }
   // the scratch holds kRedMaxGrid partials; the grid never takes more
   const int grid = RedGrid(d->NmbLin);
   if(grid > kRedMaxGrid)
   {
      printf("Reduction of %lld lines needs %d partials, the scratch holds %d\n", d->NmbLin, grid, kRedMaxGrid);
      return -4;
   }
   const int rc = ReduceLaunch(g->stream, (const float *)d->dev, d->NmbLin, RedOpp, s);
   if(rc != 1) return rc;
   if(hipMemcpyAsync(g->red_host, &s->result, sizeof(double), hipMemcpyDeviceToHost, g->stream) != hipSuccess
   || hipStreamSynchronize(g->stream) != hipSuccess)
      return -6;
   *nrm = *g->red_host;
   return 1;
}

double GmlGetReduceRunTime(size_t GmlIdx, int RedOpp)
{
   GML_CTX(g, GmlIdx);
   if(RedOpp < 0 || RedOpp >= GmlMaxRed)
   {
      printf("Invalid operation code %d\n", RedOpp);
      return -1.;
   }
   if(g->host_only) return -1.;
   HIPCHK(hipStreamSynchronize(g->stream));
   HarvestEvents(g->red_evts[RedOpp], &g->red_time[RedOpp]);
   return g->red_time[RedOpp];
}

} // extern "C"
